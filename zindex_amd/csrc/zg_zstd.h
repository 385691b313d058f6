// zg_zstd.h -- the public libzstd entry points this library links against.
// The image ships libzstd.so.1 without its development header, so the stable
// public prototypes of zstd.h (v1.4+) are declared here.
#pragma once

#include <cstddef>

extern "C" {
typedef struct ZSTD_CCtx_s ZSTD_CCtx;
typedef struct ZSTD_DCtx_s ZSTD_DCtx;
typedef struct ZSTD_inBuffer_s {
  const void *src;
  size_t size;
  size_t pos;
} ZSTD_inBuffer;
typedef struct ZSTD_outBuffer_s {
  void *dst;
  size_t size;
  size_t pos;
} ZSTD_outBuffer;

size_t ZSTD_compressBound(size_t srcSize);
unsigned ZSTD_isError(size_t code);
const char *ZSTD_getErrorName(size_t code);
unsigned long long ZSTD_getFrameContentSize(const void *src, size_t srcSize);
size_t ZSTD_findFrameCompressedSize(const void *src, size_t srcSize);

ZSTD_CCtx *ZSTD_createCCtx(void);
size_t ZSTD_freeCCtx(ZSTD_CCtx *cctx);
size_t ZSTD_CCtx_setParameter(ZSTD_CCtx *cctx, int param, int value);
size_t ZSTD_compress2(ZSTD_CCtx *cctx, void *dst, size_t dstCapacity,
                      const void *src, size_t srcSize);
size_t ZSTD_compressStream2(ZSTD_CCtx *cctx, ZSTD_outBuffer *output,
                            ZSTD_inBuffer *input, int endOp);

ZSTD_DCtx *ZSTD_createDCtx(void);
size_t ZSTD_freeDCtx(ZSTD_DCtx *dctx);
size_t ZSTD_decompressDCtx(ZSTD_DCtx *dctx, void *dst, size_t dstCapacity,
                           const void *src, size_t srcSize);
size_t ZSTD_DCtx_reset(ZSTD_DCtx *dctx, int reset);
size_t ZSTD_decompressStream(ZSTD_DCtx *zds, ZSTD_outBuffer *output,
                             ZSTD_inBuffer *input);
size_t ZSTD_DStreamInSize(void);
size_t ZSTD_DStreamOutSize(void);
}

namespace zg {
// Values of the public zstd.h enums used here.
constexpr int kZstdCompressionLevel = 100;  // ZSTD_c_compressionLevel
constexpr int kZstdChecksumFlag = 201;      // ZSTD_c_checksumFlag
constexpr int kZstdEndContinue = 0;         // ZSTD_e_continue
constexpr int kZstdEndFlush = 1;            // ZSTD_e_flush
constexpr int kZstdEndEnd = 2;              // ZSTD_e_end
constexpr int kZstdResetSessionOnly = 1;    // ZSTD_reset_session_only
constexpr unsigned long long kZstdContentSizeUnknown = 0ULL - 1;
constexpr unsigned long long kZstdContentSizeError = 0ULL - 2;
}  // namespace zg
