/*
 * oracle.c -- CPU restatement of the zorindex query path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * The product library never links, loads or calls this file.
 *
 * Pinning (how each function is checked against something other than itself;
 * the tests live in tests/test_oracle.py):
 *   zo_splitmix64        golden values splitmix64(0..9), src/lib.rs:1152-1161
 *   zo_mixsplit          relation cases, src/lib.rs:1163-1181
 *   zo_encode_base / zo_encode_full_kmer
 *                        base and full-kmer cases, src/lib.rs:891-975
 *   zo_canonical_encoded brute-force reverse complement in the test
 *   zo_scan_kmers        count cases src/lib.rs:986-1001; independent window
 *                        folding + reverse-complement invariance
 *                        (src/lib.rs:821-868); README unique_kmers known answers
 *                        (README.md:58-63) over the bundled genomes
 *   zo_scan_ocm          brute-force zo_scan_ocm_bruteforce (the reference's
 *                        own test oracle, src/open_closed_minimizers.rs:43-85),
 *                        reverse-complement invariance and the window-local
 *                        property (src/open_closed_minimizers.rs:88-128)
 *   zo_positions         closed form of storage.rs:4800-4826, slot ranges
 *   zo_fingerprint / zo_abandoned_hash / zo_partition_id
 *                        closed forms; collision golden
 *                        hash(83843,69)==hash(95900,69) (storage.rs:5446-5449)
 *   zo_screen_rows       == materialized screen for all b, counts and H
 *                        (storage.rs:4756-4797); full-hit zero rows
 *   zo_lookup            union_index fixture known answers
 *                        (storage.rs:5232-5340, 5426-5441), collision union
 *                        (storage.rs:5460-5468), plane == legacy layout
 *                        (storage.rs:5343-5423); no false negatives and
 *                        FPR ~ 2^-b on built indexes (option_coverage_cli.rs)
 *   zo_homopolymer_compress
 *                        known answer AaAACccNNNTttAA -> ACNNNTA and the
 *                        repeated-run feature equality (minimizer_schemes_cli.rs:
 *                        248-282)
 *   zo_spaced_project / zo_spaced_canonical
 *                        projection known answer, equal canonical keys and
 *                        reverse-complement equality (src/anchors.rs:750-773);
 *                        all-retained pattern leaves keys unchanged
 *                        (src/anchors.rs:839-852); bit-run extraction
 *   zo_read_features     modimizer sampling-one cases (src/lib.rs:1089-1113),
 *                        nesting and density of sampled sets
 *   zo_findere_sequence  known answers of src/findere.rs:466-490 (consecutive
 *                        hits per colour, short records, N breaks), brute-force
 *                        window intersection incl. reverse complements
 *                        (src/findere.rs:493-529), brute-force k-window rule
 *                        under the synthetic hit function (src/findere.rs:280-285)
 *   zo_strobe_hash       known answer strobe_hash(0x3332316f6c6c6568) =
 *                        0x2d119fced5ee39ce (src/anchors.rs:823-838)
 *   zo_link_strobes      checksums of the upstream strobealign iterator
 *                        witnesses (src/anchors.rs:950-988); inclusive
 *                        distances and self fallback (src/anchors.rs:1037-1062)
 *   randstrobe features  break and compression equalities (src/anchors.rs:
 *                        1013-1035); two-orientation multiset invariance
 *                        (src/anchors.rs:868-908)
 *   zo_query_reads       README known answers (970/970) on bundled data;
 *                        exact k-mer membership lower bound
 *                        (option_coverage_cli.rs:137-244)
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* L0                                                                      */
/* ---------------------------------------------------------------------- */

uint64_t zo_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

uint64_t zo_mixsplit(uint64_t key, uint64_t seed) {
  return zo_splitmix64(key + seed);
}

int zo_encode_base(uint8_t base) {
  switch (base) {
  case 'A': case 'a': return 0;
  case 'C': case 'c': return 1;
  case 'G': case 'g': return 2;
  case 'T': case 't': return 3;
  default: return -1;
  }
}

int zo_encode_full_kmer(const uint8_t *seq, size_t len, uint64_t *out) {
  uint64_t v = 0;
  for (size_t i = 0; i < len; i++) {
    int c = zo_encode_base(seq[i]);
    if (c < 0) return 0;
    v = (v << 2) | (uint64_t)c;
  }
  *out = v;
  return 1;
}

uint64_t zo_canonical_encoded(uint64_t encoded, int k) {
  uint64_t value = encoded, rc = 0;
  for (int i = 0; i < k; i++) {
    rc = (rc << 2) | ((value & 3) ^ 3);
    value >>= 2;
  }
  return encoded < rc ? encoded : rc;
}

#define FINGERPRINT_SEED 0x8E2D4F6A9B13C5D7ULL
#define ABANDONED_SEED 0xD6E8FEB86659FD93ULL
#define PARTITION_SEED 0x34B61D8F29A7C053ULL

uint64_t zo_fingerprint(uint64_t key, uint64_t seed, int bits) {
  uint64_t mask = bits >= 64 ? ~0ULL : ((1ULL << bits) - 1);
  return zo_splitmix64(key ^ seed ^ FINGERPRINT_SEED) & mask;
}

uint32_t zo_abandoned_hash(uint64_t key, uint64_t seed) {
  return (uint32_t)(zo_splitmix64(key ^ seed ^ ABANDONED_SEED) >> 32);
}

uint64_t zo_partition_id(uint64_t key, uint64_t seed, uint64_t partitions) {
  return zo_splitmix64(key ^ seed ^ PARTITION_SEED) & (partitions - 1);
}

static uint64_t ilog2_u64(uint64_t v) {
  uint64_t r = 0;
  while (v >>= 1) r++;
  return r;
}

void zo_positions(uint64_t key, uint64_t seed, uint64_t cells, int hashes,
                  uint64_t segment_length, uint64_t *out) {
  uint64_t first = zo_mixsplit(key, seed);
  uint64_t step = zo_splitmix64(first);
  uint64_t length = segment_length;
  if (length == 0) {
    uint64_t n = cells / (uint64_t)hashes;
    if (n < 1) n = 1;
    if (n > 4096) n = 4096;
    length = 1ULL << ilog2_u64(n);
  }
  uint64_t count = cells / length - (uint64_t)hashes + 1;
  uint64_t base =
      (uint64_t)(((unsigned __int128)first * (unsigned __int128)count) >> 64) *
      length;
  uint64_t h = first;
  for (int i = 0; i < hashes; i++) {
    out[i] = base + (uint64_t)i * length + ((h ^ (h >> 33)) & (length - 1));
    h += step;
  }
}

uint64_t zo_get_bits(const uint8_t *row, size_t offset, size_t width) {
  size_t byte = offset / 8, shift = offset % 8;
  size_t count = (shift + width + 7) / 8;
  uint8_t buf[8] = {0};
  size_t first = count < 8 ? count : 8;
  memcpy(buf, row + byte, first);
  uint64_t v = 0;
  for (int i = 7; i >= 0; i--) v = (v << 8) | buf[i];
  v >>= shift;
  if (count == 9) v |= (uint64_t)row[byte + 8] << (64 - shift);
  if (width < 64) v &= (1ULL << width) - 1;
  return v;
}

void zo_put_bits(uint8_t *row, size_t row_len, size_t offset, size_t width,
                 uint64_t value) {
  (void)row_len;
  if (width < 64) value &= (1ULL << width) - 1;
  size_t left = width, bit = offset, consumed = 0;
  while (left > 0) {
    size_t chunk = 8 - bit % 8;
    if (chunk > left) chunk = left;
    uint64_t mask = (1ULL << chunk) - 1;
    row[bit / 8] |= (uint8_t)(((value >> consumed) & mask) << (bit % 8));
    left -= chunk;
    bit += chunk;
    consumed += chunk;
  }
}

/* ---------------------------------------------------------------------- */
/* L1 seed extraction                                                      */
/* ---------------------------------------------------------------------- */

#define KMER_HASH_BASE 0x9e3779b185ebca87ULL

static size_t scan_long_kmers(const uint8_t *seq, size_t len, int k,
                              uint64_t *out) {
  /* src/lib.rs:687-731 */
  uint64_t inverse = KMER_HASH_BASE;
  for (int i = 0; i < 6; i++)
    inverse = inverse * (2ULL - KMER_HASH_BASE * inverse);
  uint64_t power = 1, base = KMER_HASH_BASE;
  uint64_t exponent = (uint64_t)k - 1;
  while (exponent) {
    if (exponent & 1) power *= base;
    base *= base;
    exponent >>= 1;
  }
  uint64_t domain = zo_splitmix64((uint64_t)k ^ 0x6b6d65722d763031ULL);
  uint64_t fwd = 0, rev = 0, weight = 1;
  size_t seen = 0, n = 0;
  for (size_t p = 0; p < len; p++) {
    int code = zo_encode_base(seq[p]);
    if (code < 0) {
      fwd = rev = 0;
      weight = 1;
      seen = 0;
      continue;
    }
    uint64_t value = (uint64_t)code + 1, comp = 4 - (uint64_t)code;
    if (seen < (size_t)k) {
      fwd = fwd * KMER_HASH_BASE + value;
      rev = rev + comp * weight;
      weight *= KMER_HASH_BASE;
      seen++;
    } else {
      uint64_t outgoing = (uint64_t)zo_encode_base(seq[p - k]);
      fwd = (fwd - (outgoing + 1) * power) * KMER_HASH_BASE + value;
      rev = (rev - (4 - outgoing)) * inverse + comp * power;
    }
    if (seen == (size_t)k) {
      uint64_t c = fwd < rev ? fwd : rev;
      if (out) out[n] = zo_splitmix64(c ^ domain);
      n++;
    }
  }
  return n;
}

size_t zo_scan_kmers(const uint8_t *seq, size_t len, int k, uint64_t *out) {
  if (k <= 0 || len < (size_t)k) return 0;
  if (k > 32) return scan_long_kmers(seq, len, k, out);
  uint64_t mask = k == 32 ? ~0ULL : ((1ULL << (2 * k)) - 1);
  int rc_shift = 2 * (k - 1);
  uint64_t fwd = 0, rev = 0;
  size_t seen = 0, n = 0;
  for (size_t p = 0; p < len; p++) {
    int code = zo_encode_base(seq[p]);
    if (code < 0) {
      fwd = rev = 0;
      seen = 0;
      continue;
    }
    fwd = ((fwd << 2) | (uint64_t)code) & mask;
    rev = (rev >> 2) | ((uint64_t)(code ^ 3) << rc_shift);
    seen++;
    if (seen >= (size_t)k) {
      if (out) out[n] = fwd < rev ? fwd : rev;
      n++;
    }
  }
  return n;
}

/* utils/ocm-minimizers/src/implementation.rs:89-98 */
static void ocm_parameters(int k, int m, int *w, int *t, int *s) {
  *w = k - m + 1;
  int r = m < 4 ? m : 4;
  *t = r + (m - r) % *w;
  *s = *t < 4 ? *t : 4;
  if (((*t - *s) & 1) != 0) *s -= 1;
}

typedef struct {
  uint8_t cls;
  uint64_t hash;
  size_t left, right;
} ocm_rank;

/* ScalarRank::combine, implementation.rs:308-318 (a is temporally left). */
static ocm_rank rank_combine(ocm_rank a, ocm_rank b) {
  if (a.cls < b.cls || (a.cls == b.cls && a.hash < b.hash)) return a;
  if (a.cls > b.cls || (a.cls == b.cls && a.hash > b.hash)) return b;
  a.right = b.right;
  return a;
}

/* scan_run_portable, implementation.rs:163-236, on one all-ACGT run. */
static size_t ocm_scan_run(const uint8_t *seq, size_t len, int k, int m,
                           uint64_t seed, size_t origin, uint64_t *pos_out,
                           uint64_t *key_out, size_t n_out) {
  uint64_t mmask = (m >= 32) ? ~0ULL : ((1ULL << (2 * m)) - 1);
  int mshift = 2 * (m - 1);
  if (k == m) {
    uint64_t f = 0, r = 0;
    for (size_t end = 0; end < len; end++) {
      uint64_t c = (uint64_t)zo_encode_base(seq[end]);
      f = ((f << 2) | c) & mmask;
      r = (r >> 2) | ((c ^ 3) << mshift);
      if (end + 1 >= (size_t)m) {
        if (pos_out) pos_out[n_out] = origin + end + 1 - m;
        if (key_out) key_out[n_out] = f < r ? f : r;
        n_out++;
      }
    }
    return n_out;
  }
  int w, t, s;
  ocm_parameters(k, m, &w, &t, &s);
  size_t inner_length = (size_t)(t - s + 1);
  size_t anchor_length = (size_t)(k - t + 1);
  uint64_t table[256];
  for (uint64_t x = 0; x < (1ULL << (2 * s)); x++)
    table[x] = zo_mixsplit(zo_canonical_encoded(x, s),
                           seed ^ 0xC6BC279692B5CC83ULL);
  uint64_t t_mask = (1ULL << (2 * t)) - 1;
  uint64_t s_mask = (1ULL << (2 * s)) - 1;
  size_t ring = 1;
  while (ring < (size_t)k) ring <<= 1;
  size_t mask = ring - 1;
  uint64_t *keys = calloc(ring, sizeof(uint64_t));
  uint64_t *inner = calloc(ring, sizeof(uint64_t));
  /* ScalarWindow */
  uint64_t *win = malloc(inner_length * sizeof(uint64_t));
  for (size_t i = 0; i < inner_length; i++) win[i] = ~0ULL;
  size_t win_index = 0;
  uint64_t win_prefix = ~0ULL;
  /* ScalarAnchors */
  ocm_rank empty = {3, ~0ULL, 0, 0};
  ocm_rank *anc = malloc(anchor_length * sizeof(ocm_rank));
  for (size_t i = 0; i < anchor_length; i++) anc[i] = empty;
  size_t anc_index = 0;
  ocm_rank anc_prefix = empty;

  uint64_t fwd = 0, rev = 0;
  size_t gt = 0;
  int have_previous = 0;
  size_t previous = 0;
  for (size_t end = 0; end < len; end++) {
    uint64_t code = (uint64_t)zo_encode_base(seq[end]);
    gt += code >= 2;
    if (end >= (size_t)k) gt -= zo_encode_base(seq[end - k]) >= 2;
    fwd = ((fwd << 2) | code) & mmask;
    rev = (rev >> 2) | ((code ^ 3) << mshift);
    uint64_t key = fwd < rev ? fwd : rev;
    if (end + 1 >= (size_t)m) keys[(end + 1 - m) & mask] = key;
    if (end + 1 < (size_t)s) continue;
    uint64_t h = table[fwd & s_mask];
    size_t inner_pos = end + 1 - s;
    inner[inner_pos & mask] = h;
    /* ScalarWindow::push */
    win[win_index] = h;
    if (h < win_prefix) win_prefix = h;
    win_index++;
    if (win_index == inner_length) {
      win_index = 0;
      for (size_t i = inner_length - 1; i-- > 0;)
        if (win[i + 1] < win[i]) win[i] = win[i + 1];
      win_prefix = h;
    }
    uint64_t inner_min = win_prefix < win[win_index] ? win_prefix : win[win_index];
    if (end + 1 < (size_t)t) continue;
    size_t pos = end + 1 - t;
    uint8_t cls;
    if (inner[(pos + (size_t)(t - s) / 2) & mask] == inner_min)
      cls = 0;
    else if (inner[pos & mask] == inner_min || h == inner_min)
      cls = 1;
    else
      cls = 2;
    uint64_t fwd_t = fwd & t_mask, rev_t = rev >> (2 * (m - t));
    uint64_t anchor_key = fwd_t < rev_t ? fwd_t : rev_t;
    ocm_rank value = {cls, zo_mixsplit(anchor_key, seed), pos, pos};
    /* ScalarAnchors::push */
    anc[anc_index] = value;
    anc_prefix = rank_combine(anc_prefix, value);
    anc_index++;
    if (anc_index == anchor_length) {
      anc_index = 0;
      for (size_t i = anchor_length - 1; i-- > 0;)
        anc[i] = rank_combine(anc[i], anc[i + 1]);
      anc_prefix = value;
    }
    ocm_rank selected = rank_combine(anc[anc_index], anc_prefix);
    if (end + 1 < (size_t)k) continue;
    size_t start = end + 1 - k;
    size_t anchor = gt > (size_t)(k / 2) ? selected.right : selected.left;
    size_t p = start + (anchor - start) % (size_t)w;
    if (!have_previous || previous != p) {
      if (pos_out) pos_out[n_out] = origin + p;
      if (key_out) key_out[n_out] = keys[p & mask];
      n_out++;
      previous = p;
      have_previous = 1;
    }
  }
  free(keys);
  free(inner);
  free(win);
  free(anc);
  return n_out;
}

/* scan_positions, implementation.rs:108-128: split at invalid bytes. */
size_t zo_scan_ocm(const uint8_t *seq, size_t len, int k, int m, uint64_t seed,
                   uint64_t *pos_out, uint64_t *key_out) {
  if (k == 0 || (k & 1) == 0 || m == 0 || m > 31 || m > k) return 0;
  size_t n = 0, start = 0;
  while (start < len) {
    size_t end = start;
    while (end < len && zo_encode_base(seq[end]) >= 0) end++;
    if (end - start >= (size_t)k)
      n = ocm_scan_run(seq + start, end - start, k, m, seed, start, pos_out,
                       key_out, n);
    start = end + 1;
  }
  return n;
}

static uint64_t canonical_of(const uint8_t *p, int len) {
  uint64_t v = 0;
  zo_encode_full_kmer(p, (size_t)len, &v);
  return zo_canonical_encoded(v, len);
}

size_t zo_scan_ocm_bruteforce(const uint8_t *run, size_t len, int k, int m,
                              uint64_t seed, uint64_t *pos_out,
                              uint64_t *key_out) {
  int w, t, s;
  ocm_parameters(k, m, &w, &t, &s);
  size_t n = 0;
  int have = 0;
  size_t last = 0;
  if (len < (size_t)k) return 0;
  for (size_t start = 0; start + (size_t)k <= len; start++) {
    const uint8_t *window = run + start;
    uint8_t best_cls = 255;
    uint64_t best_hash = 0;
    size_t first_tie = 0, last_tie = 0;
    for (int pos = 0; pos + t <= k; pos++) {
      const uint8_t *anchor = window + pos;
      uint64_t hs[64];
      int nh = t - s + 1;
      uint64_t minimum = ~0ULL;
      for (int i = 0; i < nh; i++) {
        hs[i] = zo_mixsplit(canonical_of(anchor + i, s),
                            seed ^ 0xC6BC279692B5CC83ULL);
        if (hs[i] < minimum) minimum = hs[i];
      }
      uint8_t cls;
      if (hs[(t - s) / 2] == minimum)
        cls = 0;
      else if (hs[0] == minimum || hs[nh - 1] == minimum)
        cls = 1;
      else
        cls = 2;
      uint64_t h = zo_mixsplit(canonical_of(anchor, t), seed);
      if (cls < best_cls || (cls == best_cls && h < best_hash)) {
        best_cls = cls;
        best_hash = h;
        first_tie = last_tie = (size_t)pos;
      } else if (cls == best_cls && h == best_hash) {
        last_tie = (size_t)pos;
      }
    }
    size_t gt = 0;
    for (int i = 0; i < k; i++) gt += zo_encode_base(window[i]) >= 2;
    size_t anchor = gt > (size_t)(k / 2) ? last_tie : first_tie;
    size_t p = start + anchor % (size_t)w;
    if (!have || last != p) {
      if (pos_out) pos_out[n] = p;
      if (key_out) key_out[n] = canonical_of(run + p, m);
      n++;
      last = p;
      have = 1;
    }
  }
  return n;
}

/* ---------------------------------------------------------------------- */
/* L3 lookup                                                               */
/* ---------------------------------------------------------------------- */

static uint64_t row_xor_bits(const uint8_t *data, const uint64_t *slots,
                             int hashes, uint64_t row_bytes, size_t offset,
                             size_t width) {
  uint64_t v = 0;
  for (int h = 0; h < hashes; h++)
    v ^= zo_get_bits(data + slots[h] * row_bytes, offset, width);
  return v;
}

size_t zo_screen_rows(const uint8_t *const *rows, int n_rows, int bits,
                      uint64_t fp, size_t count, uint32_t *hits) {
  size_t n = 0;
  for (size_t i = 0; i < count; i++) {
    uint64_t v = 0;
    for (int r = 0; r < n_rows; r++)
      v ^= zo_get_bits(rows[r], i * (size_t)bits, (size_t)bits);
    if (v == fp) hits[n++] = (uint32_t)i;
  }
  return n;
}

size_t zo_screen_materialized(const uint8_t *row, int bits, uint64_t fp,
                              size_t count, uint32_t *hits) {
  /* storage.rs:3384-3450: each field of the already-XORed row compared. */
  size_t n = 0;
  if (bits == 1) {
    for (size_t i = 0; i < count; i++) {
      uint64_t bit = (row[i / 8] >> (i % 8)) & 1;
      if (bit == fp) hits[n++] = (uint32_t)i;
    }
    return n;
  }
  for (size_t i = 0; i < count; i++) {
    uint64_t v = 0;
    size_t bytes = (size_t)bits >= 8 ? (size_t)bits / 8 : 0;
    if (bytes) {
      for (size_t b = bytes; b-- > 0;) v = (v << 8) | row[i * bytes + b];
    } else {
      size_t bit = i * (size_t)bits;
      v = (row[bit / 8] >> (bit % 8)) & ((1u << bits) - 1);
    }
    if (v == fp) hits[n++] = (uint32_t)i;
  }
  return n;
}

static const uint32_t *recovered_origins(const zo_recovery *rec, uint32_t h,
                                         uint64_t *count) {
  *count = 0;
  if (!rec || rec->n == 0) return NULL;
  uint64_t lo = 0, hi = rec->n;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (rec->hashes[mid] < h)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < rec->n && rec->hashes[lo] == h) {
    *count = rec->off[lo + 1] - rec->off[lo];
    return rec->origins + rec->off[lo];
  }
  return NULL;
}

static void lookup_bank(const zo_index *idx, const zo_bank *bank, uint64_t key,
                        uint64_t fp, uint64_t weight, uint64_t *scores) {
  uint64_t slots[32];
  int H = idx->hashes;
  int b = idx->bits;
  zo_positions(key, idx->seed, bank->slots, H, idx->segment_length, slots);
  uint64_t row_bytes = (bank->row_bits + 7) / 8;
  uint64_t G = bank->n_groups;
  if (bank->fingerprint_first) {
    /* Byte-aligned widths: XOR the L-byte fingerprint planes in 64-bit words
     * first (the reference's scalar screen, storage.rs:3285-3307), then
     * compare each field. Narrow widths compare field by field. */
    uint64_t plane = (G * (uint64_t)b + 7) / 8;
    uint8_t stack_buf[4096];
    uint8_t *xr = NULL;
    if (b >= 8) {
      xr = plane <= sizeof(stack_buf) ? stack_buf : malloc(plane);
      memcpy(xr, bank->data + slots[0] * row_bytes, plane);
      for (int h = 1; h < H; h++) {
        const uint8_t *src = bank->data + slots[h] * row_bytes;
        uint64_t w = 0;
        for (; w + 8 <= plane; w += 8) {
          uint64_t a, c;
          memcpy(&a, xr + w, 8);
          memcpy(&c, src + w, 8);
          a ^= c;
          memcpy(xr + w, &a, 8);
        }
        for (; w < plane; w++) xr[w] ^= src[w];
      }
    }
    for (uint64_t i = 0; i < G; i++) {
      uint64_t v;
      if (xr) {
        if (b == 8)
          v = xr[i];
        else if (b == 16)
          v = (uint64_t)xr[2 * i] | ((uint64_t)xr[2 * i + 1] << 8);
        else
          v = (uint64_t)xr[4 * i] | ((uint64_t)xr[4 * i + 1] << 8) |
              ((uint64_t)xr[4 * i + 2] << 16) | ((uint64_t)xr[4 * i + 3] << 24);
      } else {
        v = row_xor_bits(bank->data, slots, H, row_bytes, i * (size_t)b,
                         (size_t)b);
      }
      if (v != fp) continue;
      uint32_t m0 = bank->member_off[i], m1 = bank->member_off[i + 1];
      if (m1 - m0 == 1) {
        scores[bank->members[m0]] += weight;
      } else {
        size_t off = G * (size_t)b + bank->offset[i] - i * (size_t)b;
        uint64_t mask = row_xor_bits(bank->data, slots, H, row_bytes, off,
                                     bank->width[i] - (size_t)b);
        while (mask) {
          int bit = __builtin_ctzll(mask);
          scores[bank->members[m0 + bit]] += weight;
          mask &= mask - 1;
        }
      }
    }
    if (xr && xr != stack_buf) free(xr);
    return;
  }
  uint64_t fp_mask = (1ULL << b) - 1;
  for (uint64_t i = 0; i < G; i++) {
    uint64_t v = row_xor_bits(bank->data, slots, H, row_bytes, bank->offset[i],
                              bank->width[i]);
    if ((v & fp_mask) != fp) continue;
    uint32_t m0 = bank->member_off[i], m1 = bank->member_off[i + 1];
    if (m1 - m0 == 1) {
      scores[bank->members[m0]] += weight;
    } else {
      uint64_t mask = v >> b;
      while (mask) {
        int bit = __builtin_ctzll(mask);
        scores[bank->members[m0 + bit]] += weight;
        mask &= mask - 1;
      }
    }
  }
}

void zo_lookup(const zo_index *idx, uint64_t key, uint64_t weight,
               uint64_t *scores) {
  uint64_t n_rec = 0;
  const uint32_t *rec = recovered_origins(
      idx->recovery, zo_abandoned_hash(key, idx->seed), &n_rec);
  uint64_t snap_small[64];
  uint64_t *snap = n_rec <= 64 ? snap_small : malloc(n_rec * sizeof(uint64_t));
  for (uint64_t i = 0; i < n_rec; i++) snap[i] = scores[rec[i]];
  uint64_t fp = zo_fingerprint(key, idx->seed, idx->bits);
  /* selected_banks, storage.rs:2717-2745 */
  uint32_t next = 0;
  while (next < idx->n_banks) {
    uint32_t start = next;
    const zo_bank *bank = &idx->banks[start];
    if (bank->partitions > 1) {
      uint64_t target = zo_partition_id(key, idx->seed, bank->partitions);
      if (bank->partition == 0 && bank->partitions <= idx->n_banks - start) {
        next += bank->partitions;
        lookup_bank(idx, &idx->banks[start + target], key, fp, weight, scores);
        continue;
      }
      next += 1;
      if (target == bank->partition)
        lookup_bank(idx, bank, key, fp, weight, scores);
    } else {
      next += 1;
      lookup_bank(idx, bank, key, fp, weight, scores);
    }
  }
  for (uint64_t i = 0; i < n_rec; i++)
    if (scores[rec[i]] == snap[i]) scores[rec[i]] += weight;
  if (snap != snap_small) free(snap);
}

size_t zo_homopolymer_compress(const uint8_t *seq, size_t len, uint8_t *out) {
  size_t n = 0;
  int prev = -1;
  for (size_t i = 0; i < len; i++) {
    int c = zo_encode_base(seq[i]);
    if (c < 0 || c != prev) out[n++] = c < 0 ? seq[i] : (uint8_t)"ACGT"[c];
    prev = c;
  }
  return n;
}

uint64_t zo_spaced_project(uint64_t encoded, uint32_t width, uint64_t mask) {
  uint64_t r = 0;
  for (int i = (int)width - 1; i >= 0; i--)
    if (mask & (3ULL << (2 * i))) r = (r << 2) | ((encoded >> (2 * i)) & 3);
  return r;
}

uint64_t zo_spaced_canonical(uint64_t encoded, uint32_t width, uint64_t mask) {
  int weight = 0;
  for (uint32_t i = 0; i < width; i++) weight += (mask >> (2 * i)) & 1;
  return zo_canonical_encoded(zo_spaced_project(encoded, width, mask), weight);
}

static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

uint64_t zo_strobe_hash(uint64_t value) {
  const uint64_t p1 = 0x9E3779B185EBCA87ULL, p2 = 0xC2B2AE3D27D4EB4FULL,
                 p3 = 0x165667B19E3779F9ULL, p4 = 0x85EBCA77C2B2AE63ULL,
                 p5 = 0x27D4EB2F165667C5ULL;
  uint64_t h = (p5 + 8) ^ (rotl(value * p2, 31) * p1);
  h = rotl(h, 27) * p1 + p4;
  h ^= h >> 33;
  h *= p2;
  h ^= h >> 29;
  h *= p3;
  return h ^ (h >> 32);
}

void zo_link_strobes(const uint64_t *pos, const uint64_t *hash, size_t n, uint64_t min_d,
                     uint64_t max_d, uint64_t *keys) {
  for (size_t i = 0; i < n; i++) {
    size_t partner = i;
    int best = 1 << 30;
    for (size_t j = i + 1; j < n && pos[j] - pos[i] <= max_d; j++) {
      if (pos[j] - pos[i] < min_d) continue;
      int rank = __builtin_popcountll((hash[i] ^ hash[j]) & 255);
      if (rank < best) {
        best = rank;
        partner = j;
        if (rank == 0) break;
      }
    }
    keys[i] = hash[i] / 2 + hash[partner] / 2;
  }
}

static int cmp_pos_pair(const void *a, const void *b) {
  const uint64_t *x = a, *y = b;
  return x[0] < y[0] ? -1 : x[0] > y[0];
}

/* anchors::scan with randstrobemers (src/anchors.rs:284-351): per run of
 * valid bases, canonical k-mers at every position or the selected minimizer
 * positions (each once), hashed after the optional pattern, then linked. */
static size_t strobe_features(const zo_index *idx, const uint8_t *seq, size_t len,
                              uint64_t *out) {
  size_t n_out = 0;
  uint64_t *pos = malloc((len + 1) * sizeof(uint64_t));
  uint64_t *key = malloc((len + 1) * sizeof(uint64_t));
  uint64_t *pairs = malloc((len + 1) * 2 * sizeof(uint64_t));
  size_t a = 0;
  while (a < len) {
    if (zo_encode_base(seq[a]) < 0) { a++; continue; }
    size_t b = a;
    while (b < len && zo_encode_base(seq[b]) >= 0) b++;
    size_t n;
    if (idx->mode == 1) {
      n = zo_scan_ocm(seq + a, b - a, idx->k, idx->m, idx->seed, pos, key);
      for (size_t i = 0; i < n; i++) { pairs[2 * i] = pos[i]; pairs[2 * i + 1] = key[i]; }
      qsort(pairs, n, 2 * sizeof(uint64_t), cmp_pos_pair);
      size_t u = 0;
      for (size_t i = 0; i < n; i++)
        if (u == 0 || pairs[2 * i] != pos[u - 1]) {
          pos[u] = pairs[2 * i];
          key[u] = pairs[2 * i + 1];
          u++;
        }
      n = u;
    } else {
      n = zo_scan_kmers(seq + a, b - a, idx->k, key);
      for (size_t i = 0; i < n; i++) pos[i] = i;
    }
    for (size_t i = 0; i < n; i++) {
      uint64_t v = idx->seed_width ? zo_spaced_canonical(key[i], idx->seed_width, idx->seed_mask)
                                   : key[i];
      key[i] = zo_strobe_hash(v);
    }
    zo_link_strobes(pos, key, n, idx->strobe_min, idx->strobe_max, out + n_out);
    n_out += n;
    a = b;
  }
  free(pos);
  free(key);
  free(pairs);
  return n_out;
}

size_t zo_read_features(const zo_index *idx, const uint8_t *seq, size_t len,
                        uint64_t *out) {
  uint8_t *compressed = NULL;
  if (idx->hpc) {
    compressed = malloc(len ? len : 1);
    len = zo_homopolymer_compress(seq, len, compressed);
    seq = compressed;
  }
  size_t n;
  if (idx->strobe_max) {
    n = strobe_features(idx, seq, len, out);
    free(compressed);
    return n;
  }
  if (idx->mode == 1) {
    n = zo_scan_ocm(seq, len, idx->k, idx->m, idx->seed, NULL, out);
  } else {
    n = zo_scan_kmers(seq, len, idx->k, out);
    if (idx->mode == 2) {
      /* scan_modimizers (src/lib.rs:796-806): canonical k-mers whose
       * mixsplit(kmer, seed) is a multiple of the sampling */
      size_t kept = 0;
      for (size_t i = 0; i < n; i++)
        if (zo_mixsplit(out[i], idx->seed) % idx->sampling == 0) out[kept++] = out[i];
      n = kept;
    }
  }
  /* anchors::scan (src/anchors.rs:265-283): the pattern masks each plain
   * feature (minimizers are selected before masking) */
  if (idx->seed_width)
    for (size_t i = 0; i < n; i++)
      out[i] = zo_spaced_canonical(out[i], idx->seed_width, idx->seed_mask);
  free(compressed);
  return n;
}

void zo_findere_sequence(const uint8_t *seq, size_t len, int k, int z, uint32_t n_colors,
                         zo_hits_fn lookup, void *ctx, uint64_t *scores, uint64_t *total,
                         uint64_t *smer_queries) {
  const int s = k - z, window = z + 1;
  uint64_t *hits = malloc((n_colors ? n_colors : 1) * sizeof(uint64_t));
  uint32_t *runs = malloc((n_colors ? n_colors : 1) * sizeof(uint32_t));
  uint64_t *keys = malloc((len ? len : 1) * sizeof(uint64_t));
  size_t a = 0;
  while (a < len) {
    if (zo_encode_base(seq[a]) < 0) { a++; continue; }
    size_t b = a;
    while (b < len && zo_encode_base(seq[b]) >= 0) b++;
    if (b - a >= (size_t)k) {
      memset(runs, 0, n_colors * sizeof(uint32_t));
      *total += (uint64_t)(b - a - k + 1);
      size_t n = zo_scan_kmers(seq + a, b - a, s, keys);
      for (size_t i = 0; i < n; i++) {
        memset(hits, 0, n_colors * sizeof(uint64_t));
        lookup(ctx, keys[i], hits);
        (*smer_queries)++;
        for (uint32_t c = 0; c < n_colors; c++) {
          runs[c] = hits[c] ? (runs[c] + 1 < (uint32_t)window ? runs[c] + 1 : (uint32_t)window) : 0;
          if (runs[c] == (uint32_t)window) scores[c]++;
        }
      }
    }
    a = b;
  }
  free(hits);
  free(runs);
  free(keys);
}

static void index_hits(void *ctx, uint64_t key, uint64_t *hits) {
  zo_lookup((const zo_index *)ctx, key, 1, hits);
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* record_query.rs:337-404: sort keys, fold equal keys into weights. */
static void lookup_sorted(const zo_index *idx, uint64_t *keys, size_t n,
                          uint64_t *scores) {
  qsort(keys, n, sizeof(uint64_t), cmp_u64);
  size_t i = 0;
  while (i < n) {
    size_t j = i;
    while (j < n && keys[j] == keys[i]) j++;
    zo_lookup(idx, keys[i], (uint64_t)(j - i), scores);
    i = j;
  }
}

typedef struct {
  const zo_index *idx;
  const uint8_t *seq;
  const uint64_t *read_off;
  const uint64_t *keys;     /* explicit keys mode */
  const uint64_t *key_off;
  uint64_t begin, end;
  uint64_t *counts;
  uint64_t *totals;
  uint64_t *agg;            /* aggregate mode: per-thread sums */
  uint64_t total;
} zo_job;

static void *query_worker(void *arg) {
  zo_job *job = arg;
  const zo_index *idx = job->idx;
  size_t cap = 0;
  uint64_t *keys = NULL;
  uint64_t *scratch = job->agg ? calloc(idx->n_colors, sizeof(uint64_t)) : NULL;
  for (uint64_t r = job->begin; r < job->end; r++) {
    size_t n = 0;
    if (job->keys) {
      n = job->key_off[r + 1] - job->key_off[r];
      if (n > cap) {
        cap = n;
        keys = realloc(keys, cap * sizeof(uint64_t));
      }
      memcpy(keys, job->keys + job->key_off[r], n * sizeof(uint64_t));
    } else if (idx->mode != 3) {
      size_t len = job->read_off[r + 1] - job->read_off[r];
      if (len > cap) {
        cap = len;
        keys = realloc(keys, cap * sizeof(uint64_t));
      }
      n = zo_read_features(idx, job->seq + job->read_off[r], len, keys);
    }
    uint64_t *scores;
    if (job->agg) {
      memset(scratch, 0, idx->n_colors * sizeof(uint64_t));
      scores = scratch;
    } else {
      scores = job->counts + r * (uint64_t)idx->n_colors;
      memset(scores, 0, idx->n_colors * sizeof(uint64_t));
    }
    if (idx->mode == 3 && !job->keys) {
      /* findere (storage.rs:3651-3700, 3770-3784): compressed first when the
       * index says so, then the positional scorer over the index lookup */
      const uint8_t *rs = job->seq + job->read_off[r];
      size_t len = job->read_off[r + 1] - job->read_off[r];
      uint8_t *compressed = NULL;
      if (idx->hpc) {
        compressed = malloc(len ? len : 1);
        len = zo_homopolymer_compress(rs, len, compressed);
        rs = compressed;
      }
      uint64_t kmers = 0, probes = 0;
      zo_findere_sequence(rs, len, idx->k, (int)idx->findere_z, idx->n_colors, index_hits,
                          (void *)idx, scores, &kmers, &probes);
      free(compressed);
      n = kmers;
    } else if (idx->strobe_max && !job->keys) {
      /* randstrobemers: both query orientations, the larger count per colour
       * (record_query.rs:649-685) */
      lookup_sorted(idx, keys, n, scores);
      size_t len = job->read_off[r + 1] - job->read_off[r];
      const uint8_t *rs = job->seq + job->read_off[r];
      uint8_t *rc = malloc(len ? len : 1);
      for (size_t i = 0; i < len; i++) {
        int c = zo_encode_base(rs[len - 1 - i]);
        rc[i] = c < 0 ? rs[len - 1 - i] : (uint8_t)"TGCA"[c];
      }
      size_t n2 = zo_read_features(idx, rc, len, keys);
      free(rc);
      uint64_t *second = calloc(idx->n_colors ? idx->n_colors : 1, sizeof(uint64_t));
      lookup_sorted(idx, keys, n2, second);
      for (uint32_t c = 0; c < idx->n_colors; c++)
        if (second[c] > scores[c]) scores[c] = second[c];
      free(second);
    } else {
      lookup_sorted(idx, keys, n, scores);
    }
    if (job->totals) job->totals[r] = n;
    job->total += n;
    if (job->agg)
      for (uint32_t c = 0; c < idx->n_colors; c++) job->agg[c] += scratch[c];
  }
  free(keys);
  free(scratch);
  return NULL;
}

static int run_jobs(const zo_index *idx, const uint8_t *seq,
                    const uint64_t *read_off, const uint64_t *keys,
                    const uint64_t *key_off, uint64_t n_reads,
                    uint64_t *counts, uint64_t *totals, uint64_t *agg,
                    uint64_t *total, int threads) {
  if (threads <= 0) threads = 1;
  if ((uint64_t)threads > n_reads) threads = n_reads ? (int)n_reads : 1;
  zo_job *jobs = calloc((size_t)threads, sizeof(zo_job));
  pthread_t *tids = calloc((size_t)threads, sizeof(pthread_t));
  uint64_t per = (n_reads + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    zo_job *j = &jobs[t];
    j->idx = idx;
    j->seq = seq;
    j->read_off = read_off;
    j->keys = keys;
    j->key_off = key_off;
    j->begin = per * t < n_reads ? per * t : n_reads;
    j->end = per * (t + 1) < n_reads ? per * (t + 1) : n_reads;
    j->counts = counts;
    j->totals = totals;
    j->agg = agg ? calloc(idx->n_colors, sizeof(uint64_t)) : NULL;
  }
  if (threads == 1) {
    query_worker(&jobs[0]);
  } else {
    for (int t = 0; t < threads; t++)
      pthread_create(&tids[t], NULL, query_worker, &jobs[t]);
    for (int t = 0; t < threads; t++) pthread_join(tids[t], NULL);
  }
  uint64_t sum = 0;
  for (int t = 0; t < threads; t++) {
    sum += jobs[t].total;
    if (agg) {
      for (uint32_t c = 0; c < idx->n_colors; c++) agg[c] += jobs[t].agg[c];
      free(jobs[t].agg);
    }
  }
  if (total) *total = sum;
  free(jobs);
  free(tids);
  return 0;
}

int zo_query_reads(const zo_index *idx, const uint8_t *seq,
                   const uint64_t *read_off, uint64_t n_reads,
                   uint64_t *counts, uint64_t *totals, int threads) {
  return run_jobs(idx, seq, read_off, NULL, NULL, n_reads, counts, totals,
                  NULL, NULL, threads);
}

uint64_t zo_query_aggregate(const zo_index *idx, const uint8_t *seq,
                            const uint64_t *read_off, uint64_t n_reads,
                            uint64_t *agg, int threads) {
  uint64_t total = 0;
  memset(agg, 0, idx->n_colors * sizeof(uint64_t));
  run_jobs(idx, seq, read_off, NULL, NULL, n_reads, NULL, NULL, agg, &total,
           threads);
  return total;
}

int zo_lookup_keys(const zo_index *idx, const uint64_t *keys,
                   const uint64_t *key_off, uint64_t n_reads,
                   uint64_t *counts, int threads) {
  return run_jobs(idx, NULL, NULL, keys, key_off, n_reads, counts, NULL, NULL,
                  NULL, threads);
}

/* ---------------------------------------------------------------------- */
/* The reference's optimised CPU lookup, restated as the timed CPU baseline */
/* ---------------------------------------------------------------------- */
/* record_query.rs:337-404 (query_feature_range): each read's keys sorted and
 * equal keys folded into weights; blocks of 16 distinct keys whose positions
 * are computed first, rows of the key 4 ahead prefetched (storage.rs:
 * 2980-3015, the start of each row stream only) while the current key is
 * looked up; byte-wide fingerprint planes screened 32 bytes at a time with
 * AVX2 (storage.rs:3309-3350). Other layouts and widths take lookup_bank.
 * Counts equal zo_lookup_keys' exactly. */
#include <immintrin.h>
#include <stdatomic.h>

#define FAST_BLOCK 16
#define FAST_AHEAD 4
#define FAST_MAX_H 16

/* LSD radix sort of 64-bit keys, 8-bit digits; a digit every key shares is
 * skipped. */
static void radix_sort_u64(uint64_t *a, uint64_t *tmp, size_t n) {
  if (n < 64) {
    for (size_t i = 1; i < n; i++) {
      uint64_t v = a[i];
      size_t j = i;
      while (j > 0 && a[j - 1] > v) {
        a[j] = a[j - 1];
        j--;
      }
      a[j] = v;
    }
    return;
  }
  uint64_t *src = a, *dst = tmp;
  for (int shift = 0; shift < 64; shift += 8) {
    size_t count[256] = {0};
    for (size_t i = 0; i < n; i++) count[(src[i] >> shift) & 255]++;
    if (count[(src[0] >> shift) & 255] == n) continue;
    size_t run = 0;
    for (int d = 0; d < 256; d++) {
      size_t c = count[d];
      count[d] = run;
      run += c;
    }
    for (size_t i = 0; i < n; i++) dst[count[(src[i] >> shift) & 255]++] = src[i];
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  if (src != a) memcpy(a, src, n * sizeof(uint64_t));
}

/* The banks a key probes (selected_banks, storage.rs:2717-2745). */
static uint32_t fast_selected(const zo_index *idx, uint64_t key, const zo_bank **out) {
  uint32_t n = 0, next = 0;
  while (next < idx->n_banks) {
    uint32_t start = next;
    const zo_bank *bank = &idx->banks[start];
    if (bank->partitions > 1) {
      uint64_t target = zo_partition_id(key, idx->seed, bank->partitions);
      if (bank->partition == 0 && bank->partitions <= idx->n_banks - start) {
        next += bank->partitions;
        out[n++] = &idx->banks[start + target];
        continue;
      }
      next += 1;
      if (target == bank->partition) out[n++] = bank;
    } else {
      next += 1;
      out[n++] = bank;
    }
  }
  return n;
}

static inline void fast_hit(const zo_index *idx, const zo_bank *bank, const uint64_t *slots,
                            uint64_t row_bytes, uint64_t i, uint64_t weight, uint64_t *scores) {
  uint32_t m0 = bank->member_off[i], m1 = bank->member_off[i + 1];
  if (m1 - m0 == 1) {
    scores[bank->members[m0]] += weight;
    return;
  }
  size_t off = bank->n_groups * (size_t)idx->bits + bank->offset[i] - i * (size_t)idx->bits;
  uint64_t mask = row_xor_bits(bank->data, slots, idx->hashes, row_bytes, off,
                               bank->width[i] - (size_t)idx->bits);
  while (mask) {
    scores[bank->members[m0 + __builtin_ctzll(mask)]] += weight;
    mask &= mask - 1;
  }
}

__attribute__((target("avx2"))) static void fast_lookup_bank(const zo_index *idx,
                                                             const zo_bank *bank,
                                                             const uint64_t *slots, uint64_t fp,
                                                             uint64_t key, uint64_t weight,
                                                             uint64_t *scores) {
  const int b = idx->bits, H = idx->hashes;
  if (!bank->fingerprint_first || (b != 8 && b != 16 && b != 32)) {
    lookup_bank(idx, bank, key, fp, weight, scores);
    return;
  }
  const uint64_t row_bytes = (bank->row_bits + 7) / 8;
  const uint64_t G = bank->n_groups, plane = G * (uint64_t)b / 8, width = (uint64_t)b / 8;
  const uint8_t *rows[FAST_MAX_H];
  for (int h = 0; h < H; h++) rows[h] = bank->data + slots[h] * row_bytes;
  const __m256i target = b == 8 ? _mm256_set1_epi8((char)fp)
                         : b == 16 ? _mm256_set1_epi16((short)fp)
                                   : _mm256_set1_epi32((int)fp);
  const uint32_t lanes = b == 8 ? 0xFFFFFFFFu : b == 16 ? 0x55555555u : 0x11111111u;
  uint64_t byte = 0;
  for (; byte + 32 <= plane; byte += 32) {
    __m256i v = _mm256_loadu_si256((const __m256i *)(rows[0] + byte));
    for (int h = 1; h < H; h++)
      v = _mm256_xor_si256(v, _mm256_loadu_si256((const __m256i *)(rows[h] + byte)));
    __m256i eq = b == 8 ? _mm256_cmpeq_epi8(v, target)
                 : b == 16 ? _mm256_cmpeq_epi16(v, target)
                           : _mm256_cmpeq_epi32(v, target);
    uint32_t mask = (uint32_t)_mm256_movemask_epi8(eq) & lanes;
    while (mask) {
      fast_hit(idx, bank, slots, row_bytes, (byte + __builtin_ctz(mask)) / width, weight, scores);
      mask &= mask - 1;
    }
  }
  for (uint64_t i = byte / width; i < G; i++) {
    uint64_t v = 0;
    for (int h = 0; h < H; h++) {
      uint64_t x = 0;
      memcpy(&x, rows[h] + i * width, width);
      v ^= x;
    }
    if (v == fp) fast_hit(idx, bank, slots, row_bytes, i, weight, scores);
  }
}

typedef struct {
  const zo_index *idx;
  const uint64_t *keys, *key_off;
  uint64_t n_reads;
  uint64_t *counts;
  atomic_uint_fast64_t *next;
} fast_job;

static void *fast_worker(void *arg) {
  fast_job *job = arg;
  const zo_index *idx = job->idx;
  const int H = idx->hashes;
  size_t cap = 0;
  uint64_t *buf = NULL, *tmp = NULL;
  const zo_bank *sel[FAST_BLOCK][64];
  uint32_t nsel[FAST_BLOCK];
  uint64_t slots[FAST_BLOCK][64 * FAST_MAX_H / 4];
  uint64_t bkeys[FAST_BLOCK], weights[FAST_BLOCK];
  uint64_t snap_small[64];
  for (;;) {
    uint64_t r0 = atomic_fetch_add(job->next, 16);
    if (r0 >= job->n_reads) break;
    uint64_t r1 = r0 + 16 < job->n_reads ? r0 + 16 : job->n_reads;
    for (uint64_t r = r0; r < r1; r++) {
      size_t n = job->key_off[r + 1] - job->key_off[r];
      uint64_t *scores = job->counts + r * (uint64_t)idx->n_colors;
      memset(scores, 0, idx->n_colors * sizeof(uint64_t));
      if (n > cap) {
        cap = n;
        buf = realloc(buf, cap * sizeof(uint64_t));
        tmp = realloc(tmp, cap * sizeof(uint64_t));
      }
      memcpy(buf, job->keys + job->key_off[r], n * sizeof(uint64_t));
      radix_sort_u64(buf, tmp, n);
      size_t i = 0;
      while (i < n) {
        /* a block of up to 16 distinct keys with their weights and rows */
        int count = 0;
        while (count < FAST_BLOCK && i < n) {
          size_t j = i;
          while (j < n && buf[j] == buf[i]) j++;
          bkeys[count] = buf[i];
          weights[count] = j - i;
          nsel[count] = fast_selected(idx, buf[i], sel[count]);
          for (uint32_t s = 0; s < nsel[count]; s++)
            zo_positions(buf[i], idx->seed, sel[count][s]->slots, H, idx->segment_length,
                         &slots[count][s * H]);
          count++;
          i = j;
        }
        for (int k = 0; k < count + FAST_AHEAD; k++) {
          if (k < count && k + FAST_AHEAD < count + FAST_AHEAD) {
            /* rows of the key FAST_AHEAD after the one looked up below */
            for (uint32_t s = 0; s < nsel[k]; s++) {
              const zo_bank *bank = sel[k][s];
              uint64_t row_bytes = (bank->row_bits + 7) / 8;
              for (int h = 0; h < H; h++)
                _mm_prefetch((const char *)(bank->data + slots[k][s * H + h] * row_bytes),
                             _MM_HINT_T0);
            }
          }
          int n_look = k - FAST_AHEAD;
          if (n_look < 0) continue;
          uint64_t key = bkeys[n_look], w = weights[n_look];
          uint64_t n_rec = 0;
          const uint32_t *rec = recovered_origins(idx->recovery,
                                                  zo_abandoned_hash(key, idx->seed), &n_rec);
          uint64_t *snap = n_rec <= 64 ? snap_small : malloc(n_rec * sizeof(uint64_t));
          for (uint64_t q = 0; q < n_rec; q++) snap[q] = scores[rec[q]];
          uint64_t fp = zo_fingerprint(key, idx->seed, idx->bits);
          for (uint32_t s = 0; s < nsel[n_look]; s++)
            fast_lookup_bank(idx, sel[n_look][s], &slots[n_look][s * H], fp, key, w, scores);
          for (uint64_t q = 0; q < n_rec; q++)
            if (scores[rec[q]] == snap[q]) scores[rec[q]] += w;
          if (snap != snap_small) free(snap);
        }
      }
    }
  }
  free(buf);
  free(tmp);
  return NULL;
}

int zo_lookup_keys_fast(const zo_index *idx, const uint64_t *keys, const uint64_t *key_off,
                        uint64_t n_reads, uint64_t *counts, int threads) {
  if (idx->hashes > FAST_MAX_H || idx->n_banks * (uint64_t)idx->hashes > 64 * FAST_MAX_H / 4 ||
      idx->n_banks > 64)
    return zo_lookup_keys(idx, keys, key_off, n_reads, counts, threads);
  if (threads <= 0) threads = 1;
  atomic_uint_fast64_t next = 0;
  fast_job job = {idx, keys, key_off, n_reads, counts, &next};
  pthread_t tids[256];
  if (threads > 256) threads = 256;
  for (int t = 0; t < threads; t++) pthread_create(&tids[t], NULL, fast_worker, &job);
  for (int t = 0; t < threads; t++) pthread_join(tids[t], NULL);
  return 0;
}
