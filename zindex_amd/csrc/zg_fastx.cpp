// zg_fastx.cpp -- FASTA/FASTQ readers with the reference's parsing semantics.
#include "zg_fastx.h"

#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "zg_format.h"
#include "zg_xzbz2.h"
#include "zg_zstd.h"

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
inline bool is_ws(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v';
}
std::string first_token(const char *b, size_t n) {
  size_t j = 0;
  while (j < n && !is_ws(b[j])) j++;
  return std::string(b, j);
}

// Lines of an in-memory text.
class TextLines : public LineSource {
 public:
  explicit TextLines(const std::string &text) : t_(text) {}
  bool next(const char *&b, size_t &n) override {
    if (i_ >= t_.size()) return false;
    size_t j = t_.find('\n', i_);
    if (j == std::string::npos) j = t_.size();
    b = t_.data() + i_;
    n = j - i_;
    i_ = j + 1;
    return true;
  }

 private:
  const std::string &t_;
  size_t i_ = 0;
};

// Decompressed bytes of a file, chunk by chunk; the format by its magic.
class ByteStream {
 public:
  explicit ByteStream(const std::string &path) : path_(path), in_(1 << 20) {
    f_ = fopen(path.c_str(), "rb");
    if (!f_) fail(-2, "opening " + path + ": " + strerror(errno));
    fill();
    const uint8_t *p = (const uint8_t *)in_.data();
    const size_t a = avail();
    if (a >= 2 && p[0] == 0x1f && p[1] == 0x8b) {
      kind_ = 1;
      memset(&z_, 0, sizeof z_);
      if (inflateInit2(&z_, 15 + 32) != Z_OK) fail(-2, "gzip: init failed");
    } else if (a >= 4 && p[0] == 0x28 && p[1] == 0xb5 && p[2] == 0x2f && p[3] == 0xfd) {
      kind_ = 2;
      zd_ = ZSTD_createDCtx();
      if (!zd_) fail(-5, "zstd: out of memory");
    } else if (a >= 6 && memcmp(p, "\xfd" "7zXZ", 5) == 0) {
      kind_ = 3;
      memset(&lz_, 0, sizeof lz_);
      if (lzma_stream_decoder(&lz_, UINT64_MAX, kLzmaConcatenated) != kLzmaOk)
        fail(-2, "xz: init failed");
    } else if (a >= 4 && memcmp(p, "BZh", 3) == 0 && p[3] >= '1' && p[3] <= '9') {
      kind_ = 4;
    }
  }
  ~ByteStream() {
    if (kind_ == 1) inflateEnd(&z_);
    if (zd_) ZSTD_freeDCtx(zd_);
    if (kind_ == 3) lzma_end(&lz_);
    if (bz_open_) BZ2_bzDecompressEnd(&bz_);
    if (f_) fclose(f_);
  }
  // Up to cap decompressed bytes into buf; 0 at the end of the input.
  size_t read(char *buf, size_t cap) {
    switch (kind_) {
      case 0: {
        if (!avail()) fill();
        const size_t n = std::min(cap, avail());
        memcpy(buf, in_.data() + pos_, n);
        pos_ += n;
        return n;
      }
      case 1: return gzip(buf, cap);
      case 2: return zstd(buf, cap);
      case 3: return xz(buf, cap);
      default: return bzip2(buf, cap);
    }
  }

 private:
  size_t avail() const { return len_ - pos_; }
  // More input once the buffer is consumed; false at the end of the file.
  bool fill() {
    if (avail()) return true;
    if (eof_) return false;
    len_ = fread(in_.data(), 1, in_.size(), f_);
    pos_ = 0;
    if (len_ < in_.size()) {
      if (ferror(f_)) fail(-2, "reading " + path_);
      eof_ = true;
    }
    return len_ > 0;
  }
  size_t gzip(char *buf, size_t cap) {
    if (done_) return 0;
    z_.next_out = (Bytef *)buf;
    z_.avail_out = (uInt)cap;
    while (z_.avail_out == (uInt)cap) {
      if (!avail() && !fill()) fail(-2, "gzip: truncated stream in " + path_);
      z_.next_in = (Bytef *)in_.data() + pos_;
      z_.avail_in = (uInt)avail();
      const int rc = inflate(&z_, Z_NO_FLUSH);
      pos_ = len_ - z_.avail_in;
      if (rc == Z_STREAM_END) {
        if (!avail() && !fill()) {
          done_ = true;
          break;
        }
        if (inflateReset(&z_) != Z_OK) fail(-2, "gzip: corrupt stream in " + path_);
        continue;  // the next gzip member
      }
      if (rc != Z_OK && rc != Z_BUF_ERROR) fail(-2, "gzip: corrupt stream in " + path_);
    }
    return cap - z_.avail_out;
  }
  size_t zstd(char *buf, size_t cap) {
    ZSTD_outBuffer o{buf, cap, 0};
    while (o.pos == 0) {
      if (!avail() && !fill()) {
        if (last_ != 0) fail(-2, "zstd: truncated stream");
        // flush what the decoder still holds
        ZSTD_inBuffer none{nullptr, 0, 0};
        const size_t r = ZSTD_decompressStream(zd_, &o, &none);
        if (ZSTD_isError(r)) fail(-2, std::string("zstd: ") + ZSTD_getErrorName(r));
        return o.pos;
      }
      ZSTD_inBuffer i{in_.data() + pos_, avail(), 0};
      const size_t r = ZSTD_decompressStream(zd_, &o, &i);
      if (ZSTD_isError(r)) fail(-2, std::string("zstd: ") + ZSTD_getErrorName(r));
      pos_ += i.pos;
      last_ = r;
    }
    return o.pos;
  }
  size_t xz(char *buf, size_t cap) {
    if (done_) return 0;
    lz_.next_out = (uint8_t *)buf;
    lz_.avail_out = cap;
    while (lz_.avail_out == cap) {
      const bool more = avail() || fill();
      lz_.next_in = (const uint8_t *)in_.data() + pos_;
      lz_.avail_in = avail();
      const int rc = lzma_code(&lz_, more ? kLzmaRun : kLzmaFinish);
      pos_ = len_ - lz_.avail_in;
      if (rc == kLzmaStreamEnd) {
        done_ = true;
        break;
      }
      if (rc != kLzmaOk || (!more && lz_.avail_out == cap))
        fail(-2, "xz: corrupt or truncated stream in " + path_);
    }
    return cap - lz_.avail_out;
  }
  size_t bzip2(char *buf, size_t cap) {
    size_t got = 0;
    while (got == 0) {
      if (!bz_open_) {
        if (!avail() && !fill()) return 0;  // no further stream
        memset(&bz_, 0, sizeof bz_);
        if (BZ2_bzDecompressInit(&bz_, 0, 0) != kBzOk) fail(-2, "bzip2: init failed");
        bz_open_ = true;
      }
      if (!avail() && !fill()) fail(-2, "bzip2: corrupt or truncated stream in " + path_);
      bz_.next_in = in_.data() + pos_;
      bz_.avail_in = (unsigned)std::min<size_t>(avail(), 1u << 30);
      bz_.next_out = buf;
      bz_.avail_out = (unsigned)std::min<size_t>(cap, 1u << 30);
      const unsigned before = bz_.avail_out;
      const int rc = BZ2_bzDecompress(&bz_);
      pos_ = (size_t)(bz_.next_in - in_.data());
      got = before - bz_.avail_out;
      if (rc == kBzStreamEnd) {
        BZ2_bzDecompressEnd(&bz_);
        bz_open_ = false;  // a further stream may follow
      } else if (rc != kBzOk) {
        fail(-2, "bzip2: corrupt or truncated stream in " + path_);
      }
    }
    return got;
  }

  std::string path_;
  FILE *f_ = nullptr;
  std::vector<char> in_;
  size_t pos_ = 0, len_ = 0;
  bool eof_ = false, done_ = false;
  int kind_ = 0;  // 0 plain, 1 gzip, 2 zstd, 3 xz, 4 bzip2
  z_stream z_;
  ZSTD_DCtx *zd_ = nullptr;
  size_t last_ = 0;
  zg_lzma_stream lz_;
  zg_bz_stream bz_;
  bool bz_open_ = false;
};

// Lines of a decompressed file, read 1 MiB at a time.
class StreamLines : public LineSource {
 public:
  explicit StreamLines(const std::string &path) : bytes_(path) {}
  bool next(const char *&b, size_t &n) override {
    for (;;) {
      const size_t j = buf_.find('\n', at_);
      if (j != std::string::npos) {
        b = buf_.data() + at_;
        n = j - at_;
        at_ = j + 1;
        return true;
      }
      if (end_) {
        if (at_ >= buf_.size()) return false;
        b = buf_.data() + at_;
        n = buf_.size() - at_;
        at_ = buf_.size();
        return true;
      }
      buf_.erase(0, at_);  // keep the partial line, read more
      at_ = 0;
      const size_t old = buf_.size();
      buf_.resize(old + kChunk);
      const size_t got = bytes_.read(&buf_[old], kChunk);
      buf_.resize(old + got);
      if (!got) end_ = true;
    }
  }

 private:
  static constexpr size_t kChunk = 1 << 20;
  ByteStream bytes_;
  std::string buf_;
  size_t at_ = 0;
  bool end_ = false;
};
}  // namespace

std::unique_ptr<LineSource> open_lines(const std::string &path) {
  return std::make_unique<StreamLines>(path);
}

bool RecordReader::next(Records &r, size_t max) {
  r = Records{};
  const char *b;
  size_t n;
  bool open = false;
  if (pending_) {
    r.names.push_back(strict_ ? pending_id_ : std::string());
    open = true;
    pending_ = false;
  }
  while (!(fastq_ == 1 && r.size() >= max) && src_.next(b, n)) {
    if (strict_) {
      while (n && is_ws(b[n - 1])) n--;
      if (!n) continue;
    } else {
      while (n && b[n - 1] == '\r') n--;
      if (fastq_ < 0) {
        if (!n) continue;
        fastq_ = b[0] == '@';
        if (!fastq_) ensure(b[0] == '>', "invalid FASTA/FASTQ input", -1);
      }
    }
    if (fastq_ == 1 || (strict_ && fastq_ < 0 && b[0] == '@')) {
      fastq_ = 1;
      if (!strict_ && !n) continue;
      std::string id;
      if (strict_) {
        ensure(b[0] == '@', "expected FASTQ identifier", -1);
        id = first_token(b + 1, n - 1);
        ensure(!id.empty(), "empty FASTQ identifier", -1);
      } else {
        ensure(b[0] == '@', "invalid FASTQ record", -1);
      }
      const size_t seq_start = r.seq.size();
      while (true) {
        ensure(src_.next(b, n), "truncated FASTQ sequence", -1);
        if (strict_) {
          while (n && (b[n - 1] == '\n' || b[n - 1] == '\r')) n--;
        } else {
          while (n && b[n - 1] == '\r') n--;
        }
        if (n && b[0] == '+') break;
        if (strict_)
          for (size_t j = 0; j < n; j++)
            ensure(!is_ws(b[j]), "whitespace inside FASTQ sequence", -1);
        r.seq.append(b, n);
      }
      const size_t len = r.seq.size() - seq_start;
      size_t qual = 0;
      while (qual < len) {
        ensure(src_.next(b, n), "truncated FASTQ quality", -1);
        if (strict_) {
          while (n && (b[n - 1] == '\n' || b[n - 1] == '\r')) n--;
        } else {
          while (n && b[n - 1] == '\r') n--;
        }
        qual += n;
        if (strict_) ensure(qual <= len, "FASTQ sequence and quality lengths differ", -1);
      }
      if (!strict_) ensure(qual == len, "FASTQ sequence and quality lengths differ", -1);
      r.names.push_back(id);
      r.off.push_back(r.seq.size());
      continue;
    }
    fastq_ = 0;
    if (n && b[0] == '>') {
      std::string id;
      if (strict_) {
        id = first_token(b + 1, n - 1);
        ensure(!id.empty(), "empty FASTA identifier", -1);
      }
      if (open) {
        r.off.push_back(r.seq.size());
        if (r.size() >= max) {  // this header opens the next batch
          pending_ = true;
          pending_id_ = id;
          records_ += r.size();
          return true;
        }
      }
      r.names.push_back(id);
      open = true;
    } else {
      if (strict_) {
        ensure(open, "--records requires FASTA or FASTQ", -1);
        for (size_t j = 0; j < n; j++)
          ensure(!is_ws(b[j]), "whitespace inside FASTA sequence", -1);
      }
      r.seq.append(b, n);
    }
  }
  if (open) r.off.push_back(r.seq.size());
  records_ += r.size();
  return r.size() > 0;
}

Records parse_records(const std::string &text) {
  TextLines lines(text);
  RecordReader reader(lines, true);
  Records r;
  reader.next(r, SIZE_MAX);
  ensure(r.size() > 0, "no FASTA/FASTQ records", -1);
  return r;
}

Records parse_sequences(const std::string &text) {
  TextLines lines(text);
  RecordReader reader(lines, false);
  Records r;
  reader.next(r, SIZE_MAX);
  return r;
}

}  // namespace zg
