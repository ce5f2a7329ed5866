/*
 * mans_oracle.c -- CPU restatement of the multi-byte ANS codec path.
 *
 * TEST INFRASTRUCTURE ONLY (see mans_oracle.h).  Reference citations are
 * relative to the reference repository root.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Floating
 * point appears in exactly one place (mo_normalize's first quantisation)
 * and must be IEEE binary32 with round-to-nearest, no contraction, as on the
 * GPU the reference runs on.
 */
#include "mans_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MO_LANES_MAX 64

static uint64_t mo_div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static uint64_t mo_round_up(uint64_t a, uint64_t b) { return mo_div_up(a, b) * b; }

static void put_u16(uint8_t* p, uint32_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static void put_u32(uint8_t* p, uint32_t v) {
  p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}
static void put_u64(uint8_t* p, uint64_t v) { put_u32(p, (uint32_t)v); put_u32(p + 4, (uint32_t)(v >> 32)); }
static uint32_t get_u16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
static uint32_t get_u32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint64_t get_u64(const uint8_t* p) { return (uint64_t)get_u32(p) | ((uint64_t)get_u32(p + 4) << 32); }

mo_format mo_format_ref(void) { mo_format f = {32u, 4096u, 1u}; return f; }
mo_format mo_format_native(void) { mo_format f = {64u, 8192u, 2u}; return f; }

/* ------------------------------------------------------------------ A1 */
/* GpuANSStatistics.h:23-136: one 256-bin count over the raw bytes; the
 * uint4 vectorisation and head/tail split there do not change the counts. */
void mo_histogram(const uint8_t* in, uint64_t n, uint32_t stride, uint32_t* hist) {
  memset(hist, 0, sizeof(uint32_t) * MO_NUM_SYMBOLS);
  for (uint64_t i = 0; i < n; ++i) hist[in[i * stride]]++;
}

/* ------------------------------------------------------------------ A2 */
static int cmp_desc_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? 1 : (x > y ? -1 : 0);
}

/* Count leading zeros with clz(0) = 32, as CUDA's __clz. */
static uint32_t mo_clz32(uint32_t v) {
  uint32_t n = 0;
  if (v == 0) return 32;
  while (!(v & 0x80000000u)) { v <<= 1; ++n; }
  return n;
}

/* GpuANSStatistics.h:186-385 (normalizeProbabilitiesFromHistogram). */
int mo_normalize(const uint32_t* counts, uint32_t total, int prob_bits, uint32_t* table) {
  memset(table, 0, sizeof(uint32_t) * 4 * MO_NUM_SYMBOLS);
  if (total == 0) return 0; /* :202-204 returns before writing the table */
  const uint32_t W = 1u << prob_bits;
  uint32_t q[MO_NUM_SYMBOLS];
  int qsum = 0;
  for (int s = 0; s < MO_NUM_SYMBOLS; ++s) {
    /* :225  qProb = kProbWeight * ((float)count / (float)totalNum), where the
     * uint32 weight converts to float, one fp32 division and one fp32
     * multiply, each rounded to nearest, then truncation to uint32. */
    volatile float ratio = (float)counts[s] / (float)total;
    volatile float prod = (float)W * ratio;
    uint32_t v = (uint32_t)prod;
    if (counts[s] > 0 && v == 0) v = 1; /* :228 */
    q[s] = v;
    qsum += (int)v; /* :230, then blockSum at :235 */
  }
  /* :241-262 sort (q << 16) | symbol in descending order; rank r holds the
   * r-th largest key. */
  uint32_t key[MO_NUM_SYMBOLS];
  for (int s = 0; s < MO_NUM_SYMBOLS; ++s) key[s] = (q[s] << 16) | (uint32_t)s;
  qsort(key, MO_NUM_SYMBOLS, sizeof(uint32_t), cmp_desc_u32);
  uint32_t rsym[MO_NUM_SYMBOLS], rq[MO_NUM_SYMBOLS];
  for (int r = 0; r < MO_NUM_SYMBOLS; ++r) { rsym[r] = key[r] & 0xffffu; rq[r] = key[r] >> 16; }

  int diff = (int)W - qsum; /* :266 */
  if (diff > 0) {
    /* :268-283: each pass adds one to every symbol whose *id* is below
     * min(diff, 256), whether or not the symbol is present. */
    while (diff > 0) {
      int it = diff < MO_NUM_SYMBOLS ? diff : MO_NUM_SYMBOLS;
      for (int r = 0; r < MO_NUM_SYMBOLS; ++r)
        if ((int)rsym[r] < it) rq[r] += 1;
      diff -= it;
    }
  } else if (diff < 0) {
    /* :284-326: subtract one from the ranks [G - it, G) of the original
     * sorted order, G = number of weights > 1. */
    diff = -diff;
    while (diff > 0) {
      int g = 0;
      for (int r = 0; r < MO_NUM_SYMBOLS; ++r) g += rq[r] > 1;
      int it = diff < g ? diff : g;
      if (it <= 0) return -1; /* unreachable for W >= 512 (see tests) */
      int start = g - it;
      for (int r = start; r < g; ++r) rq[r] -= 1;
      diff -= it;
    }
  }
  uint32_t pdf[MO_NUM_SYMBOLS];
  for (int r = 0; r < MO_NUM_SYMBOLS; ++r) pdf[rsym[r]] = rq[r]; /* :329-336 */
  uint32_t cdf = 0;                                             /* :347-352 */
  for (int s = 0; s < MO_NUM_SYMBOLS; ++s) {
    uint32_t* e = table + 4 * s;
    e[0] = pdf[s];
    e[1] = cdf;
    cdf += pdf[s];
    if (pdf[s] > 0) {
      /* :359-375 */
      uint32_t shift = 32u - mo_clz32(pdf[s] - 1u);
      uint64_t magic64 = ((((uint64_t)1) << 32) * ((((uint64_t)1) << shift) - pdf[s])) / pdf[s] + 1u;
      e[2] = (uint32_t)magic64;
      e[3] = shift;
    } else {
      /* The reference divides by zero here; the entry is never used by
       * the encoder and never written to the stream.  Defined as 0. */
      e[2] = 0;
      e[3] = 0;
    }
  }
  return 0;
}

/* GpuANSEncode.h:54-65: t = umulhi(x, magic); q = (t + x) >> shift. */
uint32_t mo_magic_div(uint32_t x, uint32_t magic, uint32_t shift) {
  uint32_t t = (uint32_t)(((uint64_t)x * magic) >> 32);
  return (t + x) >> shift;
}

/* ------------------------------------------------------------- A3-A6 */
uint64_t mo_stream_overhead(uint64_t nb, mo_format f) {
  /* GpuANSCodec.h:69-82: header + probs + states + roundUp(nb, 2) uint2 */
  return MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS + 4u * f.lanes * nb + 8u * mo_round_up(nb, 2);
}

uint32_t mo_lane_max_words(uint32_t nsym, int prob_bits) {
  /* Each encode step multiplies the virtual state x * 2^(16 w) by at most
   * (2^p / pdf)(1 + 2^(p-15)) <= 2^(p + 0.1); the state starts at 2^15 and
   * ends >= 2^15, so 16 w <= nsym (p + 0.1).  One word of slack. */
  uint64_t num = (uint64_t)nsym * (uint64_t)(10 * prob_bits + 1);
  return (uint32_t)((num + 159u) / 160u) + 1u;
}

static uint32_t unit_max_words(mo_format f, int p) {
  return f.lanes * mo_lane_max_words(f.block_bytes / f.lanes, p);
}

uint64_t mo_stream_max_size(uint64_t n, int prob_bits, mo_format f) {
  uint64_t nb = mo_div_up(n, f.block_bytes);
  return mo_stream_overhead(nb, f) + 2u * nb * mo_round_up(unit_max_words(f, prob_bits), 8);
}

typedef struct {
  const uint8_t* src;
  uint32_t stride;
  uint64_t nsym;
  uint64_t nb;
  uint64_t section;    /* byte offset of the plane's ANS header in out */
  uint32_t table[4 * MO_NUM_SYMBOLS];
  uint64_t words_sum;  /* sum of padded words of the plane's units */
} mo_plane;

/* One block: GpuANSEncode.h:75-164 with encodeOne :24-73.  Symbol i of the
 * block goes to lane i % L at step i / L; steps run in increasing order and
 * within a step the lanes that emit append their low 16 bits in lane order
 * (ballot + popc of lower lanes, :45-52). */
static uint32_t enc_unit(const mo_plane* pl, uint64_t b, mo_format f, int p,
                         uint32_t* states, uint16_t* words) {
  const uint64_t start = b * f.block_bytes;
  const uint32_t n = (uint32_t)((pl->nsym - start) < f.block_bytes ? (pl->nsym - start) : f.block_bytes);
  const uint32_t L = f.lanes;
  uint32_t x[MO_LANES_MAX];
  uint32_t nw = 0;
  for (uint32_t l = 0; l < L; ++l) x[l] = MO_MIN_STATE; /* kANSStartState */
  const uint32_t steps = (uint32_t)mo_div_up(n, L);
  for (uint32_t s = 0; s < steps; ++s) {
    for (uint32_t l = 0; l < L; ++l) {
      uint32_t i = s * L + l;
      if (i >= n) break;
      uint32_t sym = pl->src[(start + i) * pl->stride];
      const uint32_t* e = pl->table + 4 * sym;
      uint32_t pdf = e[0], cdf = e[1], magic = e[2], shift = e[3];
      uint32_t check = pdf * (1u << (MO_STATE_BITS - p)); /* :40-43 */
      if (x[l] >= check) {
        words[nw++] = (uint16_t)(x[l] & 0xffffu);
        x[l] >>= MO_ENCODED_BITS;
      }
      uint32_t q = mo_magic_div(x[l], magic, shift);
      uint32_t mod = x[l] - q * pdf;
      x[l] = q * (1u << p) + mod + cdf; /* :67-69 */
    }
  }
  for (uint32_t l = 0; l < L; ++l) states[l] = x[l];
  return nw;
}

/* Encode all units of all planes.  order 0: unit u -> (plane u % P,
 * block u / P); order 1: planes concatenated.  Unit u's padded words go at
 * word offset y_u = sum_{v<u} roundUp(W_v, 8) of the data region
 * (BatchPrefixSum.h:29-104 + Align, GpuANSEncode.h:166-178, 365-378). */
static int encode_units(mo_plane* pl, int P, int order, mo_format f, int p,
                        uint8_t* out, uint64_t data_off, uint64_t* data_words,
                        int threads) {
  uint64_t total_units = 0;
  uint64_t base[5];
  for (int k = 0; k < P; ++k) { base[k] = total_units; total_units += pl[k].nb; pl[k].words_sum = 0; }
  base[P] = total_units;
  const uint32_t cap = unit_max_words(f, p);
  const uint64_t chunk = 4096;
  uint16_t* tmpw = (uint16_t*)malloc(sizeof(uint16_t) * cap * chunk);
  uint32_t* tmps = (uint32_t*)malloc(sizeof(uint32_t) * f.lanes * chunk);
  uint32_t* tmpn = (uint32_t*)malloc(sizeof(uint32_t) * chunk);
  if (!tmpw || !tmps || !tmpn) { free(tmpw); free(tmps); free(tmpn); return -3; }
  uint64_t y = 0;
  for (uint64_t c0 = 0; c0 < total_units; c0 += chunk) {
    int64_t cn = (int64_t)((total_units - c0) < chunk ? (total_units - c0) : chunk);
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : 1)
    for (int64_t j = 0; j < cn; ++j) {
      uint64_t u = c0 + (uint64_t)j;
      int k; uint64_t b;
      if (order == 0) { k = (int)(u % (uint64_t)P); b = u / (uint64_t)P; }
      else { k = 0; while (u >= base[k + 1]) ++k; b = u - base[k]; }
      tmpn[j] = enc_unit(&pl[k], b, f, p, tmps + (uint64_t)j * f.lanes, tmpw + (uint64_t)j * cap);
    }
    for (int64_t j = 0; j < cn; ++j) {
      uint64_t u = c0 + (uint64_t)j;
      int k; uint64_t b;
      if (order == 0) { k = (int)(u % (uint64_t)P); b = u / (uint64_t)P; }
      else { k = 0; while (u >= base[k + 1]) ++k; b = u - base[k]; }
      uint64_t start = b * f.block_bytes;
      uint32_t n = (uint32_t)((pl[k].nsym - start) < f.block_bytes ? (pl[k].nsym - start) : f.block_bytes);
      uint32_t w = tmpn[j];
      uint64_t padded = mo_round_up(w, 8);
      uint8_t* sec = out + pl[k].section;
      /* states: GpuANSEncode.h:157, coalesced by :248-253 */
      uint8_t* st = sec + MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS + 4u * f.lanes * b;
      for (uint32_t l = 0; l < f.lanes; ++l) put_u32(st + 4u * l, tmps[(uint64_t)j * f.lanes + l]);
      /* blockWords: GpuANSEncode.h:258-268 */
      uint8_t* bw = sec + MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS + 4u * f.lanes * pl[k].nb + 8u * b;
      put_u32(bw, (n << 16) | w);
      put_u32(bw + 4, (uint32_t)y);
      /* data: GpuANSEncode.h:270-283, padding words defined as zero */
      uint8_t* d = out + data_off + 2u * y;
      const uint16_t* src = tmpw + (uint64_t)j * cap;
      for (uint32_t i = 0; i < w; ++i) put_u16(d + 2u * i, src[i]);
      for (uint64_t i = w; i < padded; ++i) put_u16(d + 2u * i, 0);
      y += padded;
      pl[k].words_sum += padded;
    }
  }
  free(tmpw); free(tmps); free(tmpn);
  *data_words = y;
  return 0;
}

/* Plane header + probs (GpuANSEncode.h:201-237 and GpuANSCodec.h:198-208).
 * options bits 31:5, checksum, unused0/1 are uninitialised in the reference
 * and defined as zero here; so is the odd-nb blockWords pad entry. */
static void write_section_header(const mo_plane* pl, mo_format f, int p, uint8_t* out, uint32_t total_words) {
  uint8_t* h = out + pl->section;
  memset(h, 0, MO_HEADER_BYTES);
  put_u32(h + 0, (MO_MAGIC << 16) | f.version);
  put_u32(h + 4, (uint32_t)pl->nb);
  put_u32(h + 8, (uint32_t)pl->nsym);
  put_u32(h + 12, total_words);
  put_u32(h + 16, (uint32_t)p);
  for (int s = 0; s < MO_NUM_SYMBOLS; ++s) put_u16(h + MO_HEADER_BYTES + 2u * s, pl->table[4 * s]);
  if (pl->nb & 1u) {
    uint8_t* pad = h + MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS + 4u * f.lanes * pl->nb + 8u * pl->nb;
    memset(pad, 0, 8);
  }
}

static int check_params(int prob_bits, mo_format f) {
  if (prob_bits < 9 || prob_bits > 11) return -1; /* GpuANSEncode.h:348-360 */
  if (!((f.lanes == 32 && f.block_bytes == 4096 && f.version == 1) ||
        (f.lanes == 64 && f.block_bytes == 8192 && f.version == 2))) return -1;
  return 0;
}

static void plane_init(mo_plane* pl, const uint8_t* src, uint32_t stride, uint64_t nsym, int p, mo_format f) {
  uint32_t hist[MO_NUM_SYMBOLS];
  pl->src = src; pl->stride = stride; pl->nsym = nsym;
  pl->nb = mo_div_up(nsym, f.block_bytes);
  mo_histogram(src, nsym, stride, hist);
  mo_normalize(hist, (uint32_t)nsym, p, pl->table);
}

int mo_stream_encode(const uint8_t* in, uint64_t n, uint32_t stride, int prob_bits, mo_format f,
                     uint8_t* out, uint64_t out_cap, uint64_t* out_size, int threads) {
  if (check_params(prob_bits, f)) return -1;
  if (n > 0xffffffffull) return -2;
  mo_plane pl;
  plane_init(&pl, in, stride, n, prob_bits, f);
  pl.section = 0;
  uint64_t data_off = mo_stream_overhead(pl.nb, f);
  if (out_cap < mo_stream_max_size(n, prob_bits, f)) return -4;
  uint64_t words = 0;
  int rc = encode_units(&pl, 1, 0, f, prob_bits, out, data_off, &words, threads);
  if (rc) return rc;
  write_section_header(&pl, f, prob_bits, out, (uint32_t)words);
  *out_size = data_off + 2u * words; /* GpuANSCodec.h:84-87 */
  return 0;
}

/* Decode LUT (GpuANSDecode.h:27-34, 304-377): entry for slot j in symbol s's
 * bucket = (j - cdf_s) << 20 | pdf_s << 8 | s.  Returns -1 if sum != 2^p. */
static int build_lut(const uint8_t* probs, int p, uint32_t* lut) {
  uint32_t cdf = 0;
  const uint32_t W = 1u << p;
  for (int s = 0; s < MO_NUM_SYMBOLS; ++s) {
    uint32_t pdf = get_u16(probs + 2 * s);
    if (pdf > W || cdf + pdf > W) return -1;
    for (uint32_t j = 0; j < pdf; ++j) lut[cdf + j] = (j << 20) | (pdf << 8) | (uint32_t)s;
    cdf += pdf;
  }
  return cdf == W ? 0 : -1;
}

/* One block, backwards: GpuANSDecode.h:48-93 (decodeOneWarp), :95-142 (the
 * partial step first), :144-200, :260-301.  Within a step the lanes that need
 * a word read from the end of the block's words, the highest such lane taking
 * the last word (popc(vote & laneMaskGe), :83-88). */
static int dec_unit(const uint32_t* lut, int p, mo_format f, const uint8_t* states,
                    const uint8_t* words, uint32_t nw, uint32_t n, uint8_t* out, uint32_t stride) {
  const uint32_t L = f.lanes, W = 1u << p;
  uint32_t x[MO_LANES_MAX];
  for (uint32_t l = 0; l < L; ++l) x[l] = get_u32(states + 4u * l);
  uint32_t ptr = nw;
  uint32_t steps = (uint32_t)mo_div_up(n, L);
  for (uint32_t s = steps; s-- > 0;) {
    for (uint32_t l = 0; l < L; ++l) {
      uint32_t i = s * L + l;
      if (i >= n) break;
      uint32_t e = lut[x[l] & (W - 1u)];
      uint32_t pdf = (e >> 8) & 0xfffu, smc = e >> 20;
      out[(uint64_t)i * stride] = (uint8_t)(e & 0xffu);
      x[l] = pdf * (x[l] >> p) + smc;
    }
    for (uint32_t l = L; l-- > 0;) {
      uint32_t i = s * L + l;
      if (i >= n) continue;
      if (x[l] < MO_MIN_STATE) {
        if (ptr == 0) return -5;
        --ptr;
        x[l] = (x[l] << MO_ENCODED_BITS) + get_u16(words + 2u * ptr);
      }
    }
  }
  /* A valid stream consumes every word and returns each lane to the start
   * state; anything else is corruption. */
  if (ptr != 0) return -6;
  for (uint32_t l = 0; l < L; ++l) if (x[l] != MO_MIN_STATE) return -6;
  return 0;
}

typedef struct {
  uint64_t section;
  uint64_t nsym;
  uint64_t nb;
  uint8_t* dst;
  uint32_t stride;
  uint32_t lut[1u << 11];
} mo_dplane;

static int decode_units(mo_dplane* pl, int P, int order, mo_format f, int p, const uint8_t* in,
                        uint64_t c, uint64_t data_off, uint64_t data_words, int threads) {
  uint64_t total = 0, base[5];
  for (int k = 0; k < P; ++k) { base[k] = total; total += pl[k].nb; }
  base[P] = total;
  const uint32_t cap = unit_max_words(f, p);
  int err = 0;
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : 1) reduction(| : err)
  for (int64_t uu = 0; uu < (int64_t)total; ++uu) {
    uint64_t u = (uint64_t)uu;
    int k; uint64_t b;
    if (order == 0) { k = (int)(u % (uint64_t)P); b = u / (uint64_t)P; }
    else { k = 0; while (u >= base[k + 1]) ++k; b = u - base[k]; }
    const uint8_t* sec = in + pl[k].section;
    const uint8_t* st = sec + MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS + 4u * f.lanes * b;
    const uint8_t* bw = sec + MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS + 4u * f.lanes * pl[k].nb + 8u * b;
    uint32_t x = get_u32(bw), y = get_u32(bw + 4);
    uint32_t n = x >> 16, w = x & 0xffffu;
    uint64_t start = b * f.block_bytes;
    uint64_t expect = (pl[k].nsym - start) < f.block_bytes ? (pl[k].nsym - start) : f.block_bytes;
    if (n != expect || w > cap || (uint64_t)y + mo_round_up(w, 8) > data_words ||
        data_off + 2u * ((uint64_t)y + w) > c) { err |= 1; continue; }
    if (dec_unit(pl[k].lut, p, f, st, in + data_off + 2u * (uint64_t)y, w, n,
                 pl[k].dst + start * pl[k].stride, pl[k].stride)) err |= 1;
  }
  return err ? -6 : 0;
}

/* Validate a plane section header (GpuANSCodec.h:110-113, GpuANSDecode.h:
 * 220-227) and build its LUT. */
static int read_section(const uint8_t* in, uint64_t c, uint64_t section, int p, mo_format f, mo_dplane* pl) {
  if (section + MO_HEADER_BYTES + 2u * MO_NUM_SYMBOLS > c) return -5;
  const uint8_t* h = in + section;
  uint32_t mv = get_u32(h);
  if ((mv >> 16) != MO_MAGIC || (mv & 0xffffu) != f.version) return -7;
  if ((get_u32(h + 16) & 0xfu) != (uint32_t)p) return -8;
  pl->section = section;
  pl->nb = get_u32(h + 4);
  pl->nsym = get_u32(h + 8);
  if (pl->nb != mo_div_up(pl->nsym, f.block_bytes)) return -6;
  if (section + mo_stream_overhead(pl->nb, f) > c) return -5;
  if (pl->nsym > 0 && build_lut(h + MO_HEADER_BYTES, p, pl->lut)) return -6;
  return 0;
}

int mo_stream_decode(const uint8_t* in, uint64_t c, int prob_bits, uint8_t* out, uint32_t stride,
                     uint64_t out_cap, uint64_t* n_out, int threads) {
  if (prob_bits < 9 || prob_bits > 11) return -1;
  if (c < MO_HEADER_BYTES) return -5;
  uint32_t version = get_u32(in) & 0xffffu;
  mo_format f = version == 2 ? mo_format_native() : mo_format_ref();
  mo_dplane* pl = (mo_dplane*)malloc(sizeof(mo_dplane));
  if (!pl) return -3;
  int rc = read_section(in, c, 0, prob_bits, f, pl);
  if (rc) { free(pl); return rc; }
  uint64_t data_off = mo_stream_overhead(pl->nb, f);
  uint64_t words = get_u32(in + 12);
  if (data_off + 2u * words > c) { free(pl); return -5; }
  if (pl->nsym > 0 && (pl->nsym - 1) * stride + 1 > out_cap) { free(pl); return -4; }
  pl->dst = out;
  pl->stride = stride;
  rc = decode_units(pl, 1, 0, f, prob_bits, in, c, data_off, words, threads);
  *n_out = pl->nsym;
  free(pl);
  return rc;
}

/* ------------------------------------------------------------- M1-M3 */
/* mapping/cpu/mappingcpuV1.cpp constants :12-16 */
#define ADM_LANES 32
#define ADM_CHUNK 16
#define ADM_GROUP (ADM_LANES * ADM_CHUNK)
#define ADM_MOD 126
#define ADM_LANE_BYTES 32 /* cmp_chunk * max_bytes_signal_per_ele_16b */

uint64_t mo_adm_groups(uint64_t n) { return mo_div_up(n, ADM_GROUP); }

/* mappingcpuV1.cpp:33-193 (compress).  Lanes of the last group that hold no
 * element are skipped by V1's write-back loop (:146), so their bytes stay
 * zero; that is reproduced here. */
int mo_adm_map(const uint16_t* x, uint64_t n, uint16_t* centers, int32_t* offsets, uint8_t* codes,
               uint8_t* bits, uint64_t bits_cap, uint64_t* bits_len) {
  const uint64_t gsize = mo_adm_groups(n);
  int in_domain = 1;
  /* 16 elements x up to 521 bits (|d| <= 65535) fit in 1056 bytes; the
   * reference's lane buffer is 32 bytes, so L > 256 is outside its domain. */
  uint8_t (*lane_bits)[1056] = (uint8_t(*)[1056])malloc(1056u * ADM_LANES);
  uint32_t lane_len[ADM_LANES];
  int32_t* M = (int32_t*)malloc(sizeof(int32_t) * (gsize ? gsize : 1));
  if (!lane_bits || !M) { free(lane_bits); free(M); return -1; }
  /* pass 1: centers, codes and M_g (:54-109) */
  for (uint64_t g = 0; g < gsize; ++g) {
    uint64_t b0 = g * ADM_GROUP, e0 = b0 + ADM_GROUP < n ? b0 + ADM_GROUP : n;
    int sum = 0, count = 0; /* :60-66, int accumulation, integer mean */
    for (uint64_t i = b0; i < e0; ++i) { sum += x[i]; count++; }
    int center = count > 0 ? sum / count : 0;
    centers[g] = (uint16_t)center;
    int maxbytes = 0;
    for (int l = 0; l < ADM_LANES; ++l) {
      uint64_t base = b0 + (uint64_t)l * ADM_CHUNK;
      int bo = 0;
      if (base < n) {
        memset(lane_bits[l], 0, 1056);
        for (int i = 0; i < ADM_CHUNK && base + i < n; ++i) {
          int val = x[base + i];
          int diff = val > center ? val - center : center - val;
          int rem = diff % ADM_MOD;
          int len = (val == center) ? 1 : (rem == 0 ? diff / ADM_MOD : diff / ADM_MOD + 1);
          uint8_t res = (val == center) ? 1 : (uint8_t)((diff + ADM_MOD - len * ADM_MOD) * 2 + (val > center ? -1 : 0) + 1);
          codes[base + i] = res;
          lane_bits[l][bo / 8] |= (uint8_t)(1u << (7 - (bo % 8)));
          bo += len;
        }
        if (bo > 8 * ADM_LANE_BYTES) in_domain = 0;
        int lb = (bo + 7) / 8;
        if (lb > maxbytes) maxbytes = lb;
      }
      lane_len[l] = (uint32_t)bo;
    }
    M[g] = maxbytes;
    (void)lane_len;
  }
  if (!in_domain) {
    /* outside the 256-bit lane buffer the reference overflows it; there is
     * no defined bit plane to produce, only the verdict */
    free(lane_bits);
    free(M);
    *bits_len = 0;
    return 0;
  }
  /* pass 2: exclusive prefix (:127-132) */
  int64_t total = 0;
  for (uint64_t g = 0; g < gsize; ++g) { offsets[g] = (int32_t)total; total += M[g]; }
  offsets[gsize] = (int32_t)total;
  *bits_len = (uint64_t)total * ADM_LANES;
  if (*bits_len > bits_cap) { free(lane_bits); free(M); return -1; }
  memset(bits, 0, *bits_len);
  /* pass 3: per-lane padding (:111-125) and write-back (:139-159) */
  for (uint64_t g = 0; g < gsize; ++g) {
    uint64_t b0 = g * ADM_GROUP;
    int center = centers[g];
    for (int l = 0; l < ADM_LANES; ++l) {
      uint64_t base = b0 + (uint64_t)l * ADM_CHUNK;
      if (base >= n) continue;
      memset(lane_bits[0], 0, 1056);
      int bo = 0;
      for (int i = 0; i < ADM_CHUNK && base + i < n; ++i) {
        int val = x[base + i];
        int diff = val > center ? val - center : center - val;
        int rem = diff % ADM_MOD;
        int len = (val == center) ? 1 : (rem == 0 ? diff / ADM_MOD : diff / ADM_MOD + 1);
        lane_bits[0][bo / 8] |= (uint8_t)(1u << (7 - (bo % 8)));
        bo += len;
      }
      int tb = M[g] * 8;
      if (bo < tb) lane_bits[0][bo / 8] |= (bo % 8 == 0) ? 0xFF : (uint8_t)(0xFF >> (bo % 8));
      uint64_t dst = (uint64_t)offsets[g] * ADM_LANES + (uint64_t)l * (uint64_t)M[g];
      for (int i = 0; i < M[g]; ++i) bits[dst + i] = lane_bits[0][i];
    }
  }
  free(lane_bits);
  free(M);
  return in_domain;
}

/* mappingcpuV1.cpp:195-276 (decompress): parse each lane's M_g bytes MSB
 * first, a one advances the element index (stop at 16), a zero counts into
 * the current element's (uint8) signal; then d = code/2 (+126 signal) and
 * x = c - d for odd codes, c + d for even ones (uint16 arithmetic). */
void mo_adm_unmap(const uint16_t* centers, const int32_t* offsets, const uint8_t* codes,
                  const uint8_t* bits, uint64_t n, uint16_t* out) {
  const uint64_t gsize = mo_adm_groups(n);
  for (uint64_t g = 0; g < gsize; ++g) {
    int len = offsets[g + 1] - offsets[g];
    for (int l = 0; l < ADM_LANES; ++l) {
      uint64_t base = g * ADM_GROUP + (uint64_t)l * ADM_CHUNK;
      if (base >= n) continue;
      uint64_t src = (uint64_t)offsets[g] * ADM_LANES + (uint64_t)l * (uint64_t)len;
      uint8_t sig[ADM_CHUNK + 1];
      memset(sig, 0, sizeof(sig));
      int idx = -1;
      for (int ob = 0; ob < len && idx < ADM_CHUNK; ++ob) {
        uint8_t byte = bits[src + ob];
        for (int i = 7; i >= 0 && idx < ADM_CHUNK; --i) {
          if ((byte >> i) & 1) idx++;
          else if (idx >= 0) sig[idx]++;
        }
      }
      uint16_t c = centers[g];
      for (int i = 0; i < ADM_CHUNK && base + i < n; ++i) {
        uint8_t code = codes[base + i];
        int d = (code % 2 == 1) ? ((code - 1) / 2) : (code / 2);
        d += sig[i] * ADM_MOD;
        out[base + i] = (code % 2 == 1) ? (uint16_t)(c - d) : (uint16_t)(c + d);
      }
    }
  }
}

/* ------------------------------------------------------------- frame */
/* Frame header, 128 bytes, little-endian:
 *   0 u32 'MANS'  4 u16 version=1  6 u8 elem_bytes  7 u8 mode
 *   8 u8 lanes  9 u8 prob_bits  10 u8 num_planes  11 u8 unit_order
 *  12 u32 block_bytes  16 u64 num_elements  24 u64 adm_groups
 *  32 u64 meta_offset  40 u64 data_offset  48 u64 data_words
 *  56 u64 total_size  64 u64 section_offset[4]  96 u64 plane_symbols[4] */
#define MO_FRAME_MAGIC 0x534E414Du
#define MO_FRAME_HEADER 128u

static uint64_t adm_meta_bytes(uint64_t gsize) { return mo_round_up(4u * (gsize + 1) + 2u * gsize, 16); }

uint64_t mo_frame_max_size(uint64_t n, uint32_t eb, int mode, int p, mo_format f) {
  uint64_t tot = MO_FRAME_HEADER;
  uint32_t cap = unit_max_words(f, p);
  if (mode == 1) {
    uint64_t g = mo_adm_groups(n);
    uint64_t nbits = g * ADM_LANES * ADM_LANE_BYTES;
    uint64_t nb0 = mo_div_up(n, f.block_bytes), nb1 = mo_div_up(nbits, f.block_bytes);
    tot += adm_meta_bytes(g) + mo_stream_overhead(nb0, f) + mo_stream_overhead(nb1, f);
    tot += 2u * (nb0 + nb1) * mo_round_up(cap, 8);
  } else {
    uint64_t nb = mo_div_up(n, f.block_bytes);
    tot += eb * (mo_stream_overhead(nb, f) + 2u * nb * mo_round_up(cap, 8));
  }
  return tot;
}

int mo_frame_encode(const void* in, uint64_t n, uint32_t eb, int mode, int p, mo_format f,
                    uint8_t* out, uint64_t out_cap, uint64_t* out_size, int threads) {
  if (check_params(p, f)) return -1;
  if (!(eb == 1 || eb == 2 || eb == 4)) return -1;
  if (mode == 1 && eb != 2) return -1;
  if (n > 0xffffffffull) return -2;
  if (out_cap < mo_frame_max_size(n, eb, mode, p, f)) return -4;
  mo_plane* pl = (mo_plane*)calloc(4, sizeof(mo_plane));
  if (!pl) return -3;
  int P, order;
  uint64_t gsize = 0, meta_off = 0, off = MO_FRAME_HEADER;
  uint8_t* codes = NULL;
  uint8_t* bits = NULL;
  if (mode == 1) {
    gsize = mo_adm_groups(n);
    meta_off = off;
    off += adm_meta_bytes(gsize);
    int32_t* offs = (int32_t*)malloc(4u * (gsize + 1));
    uint16_t* cen = (uint16_t*)malloc(2u * (gsize ? gsize : 1));
    codes = (uint8_t*)malloc(n ? n : 1);
    uint64_t cap = gsize * ADM_LANES * ADM_LANE_BYTES, blen = 0;
    bits = (uint8_t*)malloc(cap ? cap : 1);
    int dom = mo_adm_map((const uint16_t*)in, n, cen, offs, codes, bits, cap, &blen);
    if (dom != 1) { free(offs); free(cen); free(codes); free(bits); free(pl); return -9; }
    memset(out + meta_off, 0, adm_meta_bytes(gsize));
    for (uint64_t g = 0; g <= gsize; ++g) put_u32(out + meta_off + 4u * g, (uint32_t)offs[g]);
    for (uint64_t g = 0; g < gsize; ++g) put_u16(out + meta_off + 4u * (gsize + 1) + 2u * g, cen[g]);
    free(offs); free(cen);
    P = 2; order = 1;
    plane_init(&pl[0], codes, 1, n, p, f);
    plane_init(&pl[1], bits, 1, blen, p, f);
  } else {
    P = (int)eb; order = 0;
    for (int k = 0; k < P; ++k) plane_init(&pl[k], (const uint8_t*)in + k, eb, n, p, f);
  }
  for (int k = 0; k < P; ++k) { pl[k].section = off; off += mo_stream_overhead(pl[k].nb, f); }
  uint64_t data_off = off, words = 0;
  int rc = encode_units(pl, P, order, f, p, out, data_off, &words, threads);
  if (rc == 0) {
    for (int k = 0; k < P; ++k) write_section_header(&pl[k], f, p, out, (uint32_t)pl[k].words_sum);
    uint8_t* h = out;
    memset(h, 0, MO_FRAME_HEADER);
    put_u32(h + 0, MO_FRAME_MAGIC);
    put_u16(h + 4, 1);
    h[6] = (uint8_t)eb; h[7] = (uint8_t)mode; h[8] = (uint8_t)f.lanes; h[9] = (uint8_t)p;
    h[10] = (uint8_t)P; h[11] = (uint8_t)order;
    put_u32(h + 12, f.block_bytes);
    put_u64(h + 16, n);
    put_u64(h + 24, gsize);
    put_u64(h + 32, meta_off);
    put_u64(h + 40, data_off);
    put_u64(h + 48, words);
    put_u64(h + 56, data_off + 2u * words);
    for (int k = 0; k < P; ++k) { put_u64(h + 64 + 8 * k, pl[k].section); put_u64(h + 96 + 8 * k, pl[k].nsym); }
    *out_size = data_off + 2u * words;
  }
  free(codes); free(bits); free(pl);
  return rc;
}

int mo_frame_decode(const uint8_t* in, uint64_t c, void* out, uint64_t out_cap, uint64_t* n_elems, int threads) {
  if (c < MO_FRAME_HEADER || get_u32(in) != MO_FRAME_MAGIC || get_u16(in + 4) != 1u) return -7;
  uint32_t eb = in[6], mode = in[7], lanes = in[8], p = in[9], P = in[10], order = in[11];
  mo_format f = lanes == 64 ? mo_format_native() : mo_format_ref();
  if (check_params((int)p, f) || f.block_bytes != get_u32(in + 12)) return -7;
  uint64_t n = get_u64(in + 16), gsize = get_u64(in + 24), meta = get_u64(in + 32);
  uint64_t data_off = get_u64(in + 40), words = get_u64(in + 48);
  if (data_off + 2u * words > c || P < 1 || P > 4) return -5;
  if (n * eb > out_cap) return -4;
  mo_dplane* pl = (mo_dplane*)calloc(4, sizeof(mo_dplane));
  if (!pl) return -3;
  int rc = 0;
  uint8_t *codes = NULL, *bits = NULL;
  for (uint32_t k = 0; k < P && rc == 0; ++k) {
    rc = read_section(in, c, get_u64(in + 64 + 8 * k), (int)p, f, &pl[k]);
    if (rc == 0 && pl[k].nsym != get_u64(in + 96 + 8 * k)) rc = -6;
  }
  if (rc == 0) {
    if (mode == 1) {
      codes = (uint8_t*)malloc(pl[0].nsym ? pl[0].nsym : 1);
      bits = (uint8_t*)malloc(pl[1].nsym ? pl[1].nsym : 1);
      pl[0].dst = codes; pl[0].stride = 1;
      pl[1].dst = bits; pl[1].stride = 1;
    } else {
      for (uint32_t k = 0; k < P; ++k) { pl[k].dst = (uint8_t*)out + k; pl[k].stride = eb; }
    }
    rc = decode_units(pl, (int)P, (int)order, f, (int)p, in, c, data_off, words, threads);
  }
  if (rc == 0 && mode == 1) {
    int32_t* offs = (int32_t*)malloc(4u * (gsize + 1));
    uint16_t* cen = (uint16_t*)malloc(2u * (gsize ? gsize : 1));
    for (uint64_t g = 0; g <= gsize; ++g) offs[g] = (int32_t)get_u32(in + meta + 4u * g);
    for (uint64_t g = 0; g < gsize; ++g) cen[g] = (uint16_t)get_u16(in + meta + 4u * (gsize + 1) + 2u * g);
    if ((uint64_t)offs[gsize] * ADM_LANES != pl[1].nsym) rc = -6;
    else mo_adm_unmap(cen, offs, codes, bits, n, (uint16_t*)out);
    free(offs); free(cen);
  }
  free(codes); free(bits); free(pl);
  *n_elems = n;
  return rc;
}
