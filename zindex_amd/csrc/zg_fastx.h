// zg_fastx.h -- FASTA/FASTQ input (plain, gzip, zstd) for queries and builds.
#pragma once

#include <cstdint>
#include <memory>
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

// Lines of a text: in memory, or a file decompressed chunk by chunk.
class LineSource {
 public:
  virtual ~LineSource() = default;
  // The next line without its '\n' (valid until the next call); false at
  // the end of the input.
  virtual bool next(const char *&b, size_t &n) = 0;
};

// A file read and decompressed (gzip, zstd, xz, bzip2 by their magic bytes;
// concatenated members or streams allowed) in bounded chunks, as lines.
std::unique_ptr<LineSource> open_lines(const std::string &path);

// The record readers above, resumable: next() fills `out` with up to `max`
// records and returns false once the input has none left (records.rs:58-148,
// the parser's bounded batches). strict: parse_records' checks and names;
// otherwise parse_sequences' reading.
class RecordReader {
 public:
  RecordReader(LineSource &src, bool strict) : src_(src), strict_(strict) {}
  bool next(Records &out, size_t max);
  uint64_t records() const { return records_; }

 private:
  LineSource &src_;
  bool strict_;
  int fastq_ = -1;         // -1 undecided, 0 FASTA, 1 FASTQ
  bool pending_ = false;   // a FASTA header read ahead: it opens the next batch
  std::string pending_id_;
  uint64_t records_ = 0;
};

}  // namespace zg
