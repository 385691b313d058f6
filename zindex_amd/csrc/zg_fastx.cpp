// zg_fastx.cpp -- FASTA/FASTQ readers with the reference's parsing semantics.
#include "zg_fastx.h"

#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "zg_format.h"
#include "zg_xzbz2.h"

namespace zg {

std::string read_maybe_compressed(const std::string &path) {
  std::string raw = read_file(path);
  const uint8_t *p = (const uint8_t *)raw.data();
  if (raw.size() >= 2 && p[0] == 0x1f && p[1] == 0x8b) {
    // gzip, possibly multi-member: inflate with automatic header detection.
    std::string out;
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit2(&z, 15 + 32) != Z_OK) fail(-2, "gzip: init failed");
    z.next_in = (Bytef *)raw.data();
    z.avail_in = (uInt)raw.size();
    std::vector<char> buf(1 << 20);
    int rc = Z_OK;
    while (true) {
      z.next_out = (Bytef *)buf.data();
      z.avail_out = (uInt)buf.size();
      rc = inflate(&z, Z_NO_FLUSH);
      out.append(buf.data(), buf.size() - z.avail_out);
      if (rc == Z_STREAM_END) {
        if (z.avail_in == 0) break;
        // next gzip member
        if (inflateReset(&z) != Z_OK) break;
        continue;
      }
      if (rc != Z_OK) {
        inflateEnd(&z);
        fail(-2, "gzip: corrupt stream in " + path);
      }
      if (z.avail_in == 0 && z.avail_out != 0) {
        inflateEnd(&z);
        fail(-2, "gzip: truncated stream in " + path);
      }
    }
    inflateEnd(&z);
    return out;
  }
  if (raw.size() >= 4 && p[0] == 0x28 && p[1] == 0xb5 && p[2] == 0x2f && p[3] == 0xfd)
    return zstd_decompress_all(raw);
  if (raw.size() >= 6 && memcmp(p, "\xfd" "7zXZ", 5) == 0) {
    // xz, concatenated streams allowed
    zg_lzma_stream z;
    memset(&z, 0, sizeof z);
    if (lzma_stream_decoder(&z, UINT64_MAX, kLzmaConcatenated) != kLzmaOk)
      fail(-2, "xz: init failed");
    z.next_in = p;
    z.avail_in = raw.size();
    std::string out;
    std::vector<uint8_t> buf(1 << 20);
    while (true) {
      z.next_out = buf.data();
      z.avail_out = buf.size();
      const int rc = lzma_code(&z, kLzmaFinish);
      out.append((const char *)buf.data(), buf.size() - z.avail_out);
      if (rc == kLzmaStreamEnd) break;
      if (rc != kLzmaOk) {
        lzma_end(&z);
        fail(-2, "xz: corrupt or truncated stream in " + path);
      }
    }
    lzma_end(&z);
    return out;
  }
  if (raw.size() >= 4 && memcmp(p, "BZh", 3) == 0 && p[3] >= '1' && p[3] <= '9') {
    // bzip2, possibly several concatenated streams
    std::string out;
    std::vector<char> buf(1 << 20);
    size_t at = 0;
    while (at < raw.size()) {
      zg_bz_stream z;
      memset(&z, 0, sizeof z);
      if (BZ2_bzDecompressInit(&z, 0, 0) != kBzOk) fail(-2, "bzip2: init failed");
      z.next_in = (char *)raw.data() + at;
      z.avail_in = (unsigned)std::min<size_t>(raw.size() - at, 1u << 30);
      int rc = kBzOk;
      while (rc == kBzOk) {
        z.next_out = buf.data();
        z.avail_out = (unsigned)buf.size();
        rc = BZ2_bzDecompress(&z);
        out.append(buf.data(), buf.size() - z.avail_out);
        if (rc == kBzOk && z.avail_in == 0 && z.avail_out != 0) rc = -100;  // truncated
      }
      const size_t used = (size_t)(z.next_in - (raw.data() + at));
      BZ2_bzDecompressEnd(&z);
      if (rc != kBzStreamEnd) fail(-2, "bzip2: corrupt or truncated stream in " + path);
      at += used;
    }
    return out;
  }
  return raw;
}

namespace {
struct LineReader {
  const std::string &t;
  size_t i = 0;
  explicit LineReader(const std::string &text) : t(text) {}
  // Returns false at end of input; line excludes the '\n'.
  bool next(const char *&b, size_t &n) {
    if (i >= t.size()) return false;
    size_t j = t.find('\n', i);
    if (j == std::string::npos) j = t.size();
    b = t.data() + i;
    n = j - i;
    i = j + 1;
    return true;
  }
};
inline bool is_ws(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v';
}
std::string first_token(const char *b, size_t n) {
  size_t j = 0;
  while (j < n && !is_ws(b[j])) j++;
  return std::string(b, j);
}
}  // namespace

Records parse_records(const std::string &text) {
  Records r;
  LineReader lr(text);
  const char *b;
  size_t n;
  int fastq = -1;
  bool open = false;
  while (lr.next(b, n)) {
    while (n && is_ws(b[n - 1])) n--;
    if (!n) continue;
    if (fastq == 1 || (fastq < 0 && b[0] == '@')) {
      fastq = 1;
      ensure(b[0] == '@', "expected FASTQ identifier", -1);
      std::string id = first_token(b + 1, n - 1);
      ensure(!id.empty(), "empty FASTQ identifier", -1);
      size_t seq_start = r.seq.size();
      while (true) {
        ensure(lr.next(b, n), "truncated FASTQ sequence", -1);
        while (n && (b[n - 1] == '\n' || b[n - 1] == '\r')) n--;
        if (n && b[0] == '+') break;
        for (size_t j = 0; j < n; j++)
          ensure(!is_ws(b[j]), "whitespace inside FASTQ sequence", -1);
        r.seq.append(b, n);
      }
      size_t len = r.seq.size() - seq_start, qual = 0;
      while (qual < len) {
        ensure(lr.next(b, n), "truncated FASTQ quality", -1);
        while (n && (b[n - 1] == '\n' || b[n - 1] == '\r')) n--;
        qual += n;
        ensure(qual <= len, "FASTQ sequence and quality lengths differ", -1);
      }
      r.names.push_back(id);
      r.off.push_back(r.seq.size());
      continue;
    }
    fastq = 0;
    if (b[0] == '>') {
      if (open) r.off.push_back(r.seq.size());
      std::string id = first_token(b + 1, n - 1);
      ensure(!id.empty(), "empty FASTA identifier", -1);
      r.names.push_back(id);
      open = true;
    } else {
      ensure(open, "--records requires FASTA or FASTQ", -1);
      for (size_t j = 0; j < n; j++)
        ensure(!is_ws(b[j]), "whitespace inside FASTA sequence", -1);
      r.seq.append(b, n);
    }
  }
  if (open) r.off.push_back(r.seq.size());
  ensure(r.size() > 0, "no FASTA/FASTQ records", -1);
  return r;
}

Records parse_sequences(const std::string &text) {
  Records r;
  LineReader lr(text);
  const char *b;
  size_t n;
  int fastq = -1;
  bool open = false;
  while (lr.next(b, n)) {
    while (n && (b[n - 1] == '\r')) n--;
    if (fastq < 0) {
      if (!n) continue;
      fastq = b[0] == '@';
      if (!fastq) ensure(b[0] == '>', "invalid FASTA/FASTQ input", -1);
    }
    if (fastq) {
      if (!n) continue;
      ensure(b[0] == '@', "invalid FASTQ record", -1);
      size_t seq_start = r.seq.size();
      while (true) {
        ensure(lr.next(b, n), "truncated FASTQ sequence", -1);
        while (n && b[n - 1] == '\r') n--;
        if (n && b[0] == '+') break;
        r.seq.append(b, n);
      }
      size_t len = r.seq.size() - seq_start, qual = 0;
      while (qual < len) {
        ensure(lr.next(b, n), "truncated FASTQ quality", -1);
        while (n && b[n - 1] == '\r') n--;
        qual += n;
      }
      ensure(qual == len, "FASTQ sequence and quality lengths differ", -1);
      r.names.emplace_back();
      r.off.push_back(r.seq.size());
      continue;
    }
    if (n && b[0] == '>') {
      if (open) r.off.push_back(r.seq.size());
      r.names.emplace_back();
      open = true;
    } else {
      r.seq.append(b, n);
    }
  }
  if (open) r.off.push_back(r.seq.size());
  return r;
}

}  // namespace zg
