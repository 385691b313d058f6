// zg_fastx.h -- FASTA/FASTQ input (plain, gzip, zstd) for queries and builds.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace zg {

// A batch of records: record i is seq[off[i], off[i+1]) named names[i].
struct Records {
  std::vector<std::string> names;
  std::string seq;
  std::vector<uint64_t> off{0};
  size_t size() const { return off.size() - 1; }
};

// Decompresses a file by its magic bytes (gzip, zstd, xz, bzip2; concatenated
// members or streams allowed) or returns it as is.
std::string read_maybe_compressed(const std::string &path);

// Record reader of records.rs:150-275 (strict: ids, FASTQ quality checks,
// no whitespace inside sequences; errors on zero records).
Records parse_records(const std::string &text);

// Sequence reader for aggregate queries and builds (src/lib.rs:27-33 parser
// configuration): every FASTA/FASTQ record's sequence, lines joined; names
// and quality are ignored. Bytes other than ACGT/acgt later split seed runs.
Records parse_sequences(const std::string &text);

}  // namespace zg
