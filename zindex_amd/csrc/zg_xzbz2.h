// zg_xzbz2.h -- the public liblzma and libbzip2 decoder entry points this
// library links against. The image ships liblzma.so.5 and libbz2.so.1 without
// their development headers, so the stable public declarations of lzma.h
// (xz 5.x) and bzlib.h (1.0.x) that the decoders use are given here.
#pragma once

#include <cstddef>
#include <cstdint>

extern "C" {
// lzma/base.h: lzma_stream (zero-initialised is LZMA_STREAM_INIT)
struct zg_lzma_stream {
  const uint8_t *next_in;
  size_t avail_in;
  uint64_t total_in;
  uint8_t *next_out;
  size_t avail_out;
  uint64_t total_out;
  const void *allocator;
  void *internal;
  void *reserved_ptr1, *reserved_ptr2, *reserved_ptr3, *reserved_ptr4;
  uint64_t reserved_int1, reserved_int2;
  size_t reserved_int3, reserved_int4;
  int reserved_enum1, reserved_enum2;
  unsigned char spare[256];  // room should a library build add trailing fields
};
int lzma_stream_decoder(zg_lzma_stream *strm, uint64_t memlimit, uint32_t flags);
int lzma_code(zg_lzma_stream *strm, int action);
void lzma_end(zg_lzma_stream *strm);

// bzlib.h: bz_stream
struct zg_bz_stream {
  char *next_in;
  unsigned int avail_in, total_in_lo32, total_in_hi32;
  char *next_out;
  unsigned int avail_out, total_out_lo32, total_out_hi32;
  void *state;
  void *(*bzalloc)(void *, int, int);
  void (*bzfree)(void *, void *);
  void *opaque;
  unsigned char spare[256];
};
int BZ2_bzDecompressInit(zg_bz_stream *strm, int verbosity, int small);
int BZ2_bzDecompress(zg_bz_stream *strm);
int BZ2_bzDecompressEnd(zg_bz_stream *strm);
}

namespace zg {
constexpr int kLzmaOk = 0, kLzmaStreamEnd = 1, kLzmaRun = 0, kLzmaFinish = 3;
constexpr uint32_t kLzmaConcatenated = 0x08;
constexpr int kBzOk = 0, kBzStreamEnd = 4;
}  // namespace zg
