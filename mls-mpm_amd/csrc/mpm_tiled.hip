// mpm_tiled.hip -- the sorted, LDS-staged forms of the P2G scatters and the
// periodic particle sort.
//
// Layout.  Particles are re-sorted by cell every `sort_period` steps with a
// counting sort whose key orders cells tile-major (4x4x4 tiles in Morton
// order, x fastest inside a tile).  The sorted sequence is cut into chunks of
// CH = 64*SL particles; inside a full chunk the particle of sorted rank
// q = lane*SL + slice is stored at slot chunk*CH + slice*64 + lane.  A wave
// reading 64 consecutive slots therefore loads coalesced memory while its
// lanes hold particles SL ranks apart, i.e. in different cells, so the LDS
// atomics of one instruction rarely hit the same address.  The last, partial
// chunk keeps sorted order; every slot below n holds a particle.
//
// P2G.  One block per chunk.  The block reduces its particles' base cells to
// a window (the box their 3x3x3 stencils cover), accumulates the encoded i32
// contributions of all its particles into that window in LDS (ds_add), and
// adds every non-zero window cell to the global grid with one i32 atomic.
// Integer addition is associative, so the grid is bit-identical to per-
// particle global atomics.  A chunk whose window exceeds the LDS budget
// (particles far apart, e.g. before the first sort) takes the direct
// per-particle global-atomic path instead.
#include <climits>

#include "mpm_kernels.h"
#include "mpm_ops.h"
#include "mpm_tiled.h"

namespace mpm {

constexpr int kTB = 256;  // threads per block in the tiled kernels

// Window of one chunk: origin and extent in cells; wx == 0 marks the
// fallback (window too large or empty chunk).
struct Box {
  int x0, y0, z0, wx, wy, wz;
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ uint32_t window_gidx(const Box &bx, const Params &P, uint32_t c) {
  uint32_t sxy = (uint32_t)(bx.wx * bx.wy);
  uint32_t cz = c / sxy;
  uint32_t rem = c - cz * sxy;
  uint32_t cy = rem / (uint32_t)bx.wx;
  uint32_t cx = rem - cy * (uint32_t)bx.wx;
  return ((uint32_t)(bx.z0 + (int)cz) * P.gy + (uint32_t)(bx.y0 + (int)cy)) * P.gx +
         (uint32_t)(bx.x0 + (int)cx);
}

// Sort key of the cell holding a particle: Morton rank of its 4x4x4 tile,
// then x-fastest position inside the tile.
__device__ __forceinline__ uint32_t cell_key(const Params &P, const KeyOut &ko, float x, float y,
                                             float z) {
  int cx = min(max((int)floorf(x), 0), (int)P.gx - 1);
  int cy = min(max((int)floorf(y), 0), (int)P.gy - 1);
  int cz = min(max((int)floorf(z), 0), (int)P.gz - 1);
  uint32_t tile = ((uint32_t)(cz >> 2) * ko.tdy + (uint32_t)(cy >> 2)) * ko.tdx + (uint32_t)(cx >> 2);
  return ko.tile_rank[tile] * 64u + (uint32_t)(((cz & 3) << 4) | ((cy & 3) << 2) | (cx & 3));
}
__device__ __forceinline__ void record_key(const Params &P, const KeyOut &ko, uint32_t s, float x,
                                           float y, float z) {
  uint32_t key = cell_key(P, ko, x, y, z);
  ko.keys[s] = key;
  ko.rank[s] = atomicAdd(&ko.counts[key], 1u);
}

// ---------------------------------------------------------------------------
// P2G mass, tiled.  Also writes the chunk's window for the momentum pass and,
// when ko.counts is set, the cell key and rank of every particle for the
// re-layout done by G2P.
template <int SL, bool UNIT>
__global__ __launch_bounds__(kTB) void k_p2g_mass_tiled(Params P, PState S, GState g,
                                                        int32_t *boxes, uint32_t wmax,
                                                        KeyOut ko) {
  constexpr int PER = SL / 4;  // particles per thread
  constexpr int CH = 64 * SL;
  extern __shared__ int32_t lds[];
  __shared__ int red[6][kTB / 64];
  __shared__ Box sbox;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint32_t cbase = blockIdx.x * (uint32_t)CH;

  float px[PER], py[PER], pz[PER];
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
  for (int r = 0; r < PER; r++) {
    uint32_t s = cbase + (uint32_t)((wv * PER + r) * 64 + lane);
    if (s < P.n) {
      px[r] = S.at(PX, s);
      py[r] = S.at(PY, s);
      pz[r] = S.at(PZ, s);
      int bx = (int)floorf(px[r]) - 1, by = (int)floorf(py[r]) - 1, bz = (int)floorf(pz[r]) - 1;
      if (ko.counts) record_key(P, ko, s, px[r], py[r], pz[r]);
      mn[0] = min(mn[0], bx); mx[0] = max(mx[0], bx);
      mn[1] = min(mn[1], by); mx[1] = max(mx[1], by);
      mn[2] = min(mn[2], bz); mx[2] = max(mx[2], bz);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int lo = wave_min(mn[a]), hi = wave_max(mx[a]);
    if (lane == 0) {
      red[a][wv] = lo;
      red[3 + a][wv] = hi;
    }
  }
  __syncthreads();
  if (t == 0) {
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      lo[a] = red[a][0];
      hi[a] = red[3 + a][0];
      for (int k = 1; k < kTB / 64; k++) {
        lo[a] = min(lo[a], red[a][k]);
        hi[a] = max(hi[a], red[3 + a][k]);
      }
    }
    Box b{0, 0, 0, 0, 0, 0};
    if (lo[0] <= hi[0]) {
      int wx = hi[0] - lo[0] + 3, wy = hi[1] - lo[1] + 3, wz = hi[2] - lo[2] + 3;
      if ((uint64_t)wx * (uint64_t)wy * (uint64_t)wz <= (uint64_t)wmax)
        b = Box{lo[0], lo[1], lo[2], wx, wy, wz};
    }
    sbox = b;
    int32_t *o = boxes + 6 * blockIdx.x;
    o[0] = b.x0; o[1] = b.y0; o[2] = b.z0; o[3] = b.wx; o[4] = b.wy; o[5] = b.wz;
  }
  __syncthreads();
  const Box bx = sbox;

  if (bx.wx == 0) {  // fallback: direct global atomics (WGSL_P2G_MASS as written)
#pragma unroll
    for (int r = 0; r < PER; r++) {
      uint32_t s = cbase + (uint32_t)((wv * PER + r) * 64 + lane);
      if (s >= P.n) continue;
      PW pw;
      particle_weights(px[r], py[r], pz[r], pw);
      float pm = UNIT ? 1.0f : S.at(PM, s);
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            int idx = cell_idx_checked(pw.b[0] + i, pw.b[1] + j, pw.b[2] + k, P);
            if (idx < 0) continue;
            float w = (pw.w[0][i] * pw.w[1][j]) * pw.w[2][k];
            atomicAdd(&g.m[idx], encode(UNIT ? w : w * pm));
          }
    }
    return;
  }

  const uint32_t vol = (uint32_t)(bx.wx * bx.wy * bx.wz);
  for (uint32_t c = t; c < vol; c += kTB) lds[c] = 0;
  __syncthreads();
  const int sxy = bx.wx * bx.wy;
#pragma unroll
  for (int r = 0; r < PER; r++) {
    uint32_t s = cbase + (uint32_t)((wv * PER + r) * 64 + lane);
    if (s >= P.n) continue;
    PW pw;
    particle_weights(px[r], py[r], pz[r], pw);
    float pm = UNIT ? 1.0f : S.at(PM, s);
    int lb = ((pw.b[2] - bx.z0) * bx.wy + (pw.b[1] - bx.y0)) * bx.wx + (pw.b[0] - bx.x0);
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        int row = lb + k * sxy + j * bx.wx;
        float wxy[3];
#pragma unroll
        for (int i = 0; i < 3; i++) wxy[i] = pw.w[0][i] * pw.w[1][j];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          float w = wxy[i] * pw.w[2][k];
          atomicAdd(&lds[row + i], encode(UNIT ? w : w * pm));
        }
      }
  }
  __syncthreads();
  for (uint32_t c = t; c < vol; c += kTB) {
    int32_t v = lds[c];
    if (v != 0) atomicAdd(&g.m[window_gidx(bx, P, c)], v);
  }
}

// ---------------------------------------------------------------------------
// P2G momentum, tiled: the decoded mass window and the momentum window live
// in LDS; density is gathered from LDS.
template <int SL, bool UNIT>
__global__ __launch_bounds__(kTB) void k_p2g_mom_tiled(Params P, PState S, GState g,
                                                       const int32_t *boxes, uint32_t wmax) {
  constexpr int PER = SL / 4;
  constexpr int CH = 64 * SL;
  extern __shared__ int32_t lds[];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint32_t cbase = blockIdx.x * (uint32_t)CH;
  const int32_t *bp = boxes + 6 * blockIdx.x;
  const Box bx{bp[0], bp[1], bp[2], bp[3], bp[4], bp[5]};

  if (bx.wx == 0) {  // fallback: decode on the fly, global atomics
    const uint32_t gx = P.gx, slice = P.gx * P.gy;
    for (int r = 0; r < PER; r++) {
      uint32_t s = cbase + (uint32_t)((wv * PER + r) * 64 + lane);
      if (s >= P.n) continue;
      PW pw;
      particle_weights(S.at(PX, s), S.at(PY, s), S.at(PZ, s), pw);
      uint32_t base = (uint32_t)((pw.b[2] * (int)P.gy + pw.b[1]) * (int)gx + pw.b[0]);
      p2g_mom_particle<UNIT>(
          P, S, s, pw,
          [&](int k, int j, int i) { return decode(g.m[base + k * slice + j * gx + i]); },
          [&](int k, int j, int i, int32_t ex, int32_t ey, int32_t ez) {
            uint32_t c = base + k * slice + j * gx + i;
            atomicAdd(&g.mx[c], ex);
            atomicAdd(&g.my[c], ey);
            atomicAdd(&g.mz[c], ez);
          });
    }
    return;
  }

  float *lm = (float *)lds;      // decoded mass, wmax
  int32_t *lv = lds + wmax;      // momentum, 3 interleaved per cell
  const uint32_t vol = (uint32_t)(bx.wx * bx.wy * bx.wz);
  for (uint32_t c = t; c < vol; c += kTB) {
    lm[c] = decode(g.m[window_gidx(bx, P, c)]);
    lv[3 * c] = 0;
    lv[3 * c + 1] = 0;
    lv[3 * c + 2] = 0;
  }
  __syncthreads();
  const int sxy = bx.wx * bx.wy;
  for (int r = 0; r < PER; r++) {
    uint32_t s = cbase + (uint32_t)((wv * PER + r) * 64 + lane);
    if (s >= P.n) continue;
    PW pw;
    particle_weights(S.at(PX, s), S.at(PY, s), S.at(PZ, s), pw);
    int lb = ((pw.b[2] - bx.z0) * bx.wy + (pw.b[1] - bx.y0)) * bx.wx + (pw.b[0] - bx.x0);
    p2g_mom_particle<UNIT>(
        P, S, s, pw, [&](int k, int j, int i) { return lm[lb + k * sxy + j * bx.wx + i]; },
        [&](int k, int j, int i, int32_t ex, int32_t ey, int32_t ez) {
          int c = lb + k * sxy + j * bx.wx + i;
          atomicAdd(&lv[3 * c], ex);
          atomicAdd(&lv[3 * c + 1], ey);
          atomicAdd(&lv[3 * c + 2], ez);
        });
  }
  __syncthreads();
  for (uint32_t c = t; c < vol; c += kTB) {
    int32_t ex = lv[3 * c], ey = lv[3 * c + 1], ez = lv[3 * c + 2];
    if ((ex | ey | ez) == 0) continue;
    uint32_t gi = window_gidx(bx, P, c);
    if (ex) atomicAdd(&g.mx[gi], ex);
    if (ey) atomicAdd(&g.my[gi], ey);
    if (ez) atomicAdd(&g.mz[gi], ez);
  }
}

// ---------------------------------------------------------------------------
// Counting sort by cell key.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *a, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_sort_keys(Params P, PState S, KeyOut ko) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.n) return;
  record_key(P, ko, s, S.at(PX, s), S.at(PY, s), S.at(PZ, s));
}

// Slot of sorted rank q: inside a full chunk of 64*sl particles, rank
// lane*sl + slice goes to slice*64 + lane; the last partial chunk keeps order.
__device__ __forceinline__ uint32_t slot_of_rank(uint32_t q, uint32_t n, uint32_t sl) {
  const uint32_t ch = 64u * sl;
  uint32_t full = (n / ch) * ch;
  if (q >= full) return q;
  uint32_t r = q % ch;
  return (q - r) + (r % sl) * 64u + r / sl;
}

// Inverse permutation for the G2P re-layout: src_of[new slot] = old slot.
__global__ __launch_bounds__(256) void k_build_src(uint32_t n, uint32_t sl, const uint32_t *keys,
                                                   const uint32_t *rank, const uint32_t *offsets,
                                                   uint32_t *src_of) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  src_of[slot_of_rank(offsets[keys[s]] + rank[s], n, sl)] = s;
}

__global__ __launch_bounds__(256) void k_sort_scatter(uint32_t n, uint32_t sl, PState src,
                                                      PState dst, const uint32_t *keys,
                                                      const uint32_t *rank,
                                                      const uint32_t *offsets) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t d = slot_of_rank(offsets[keys[s]] + rank[s], n, sl);
#pragma unroll
  for (int k = 0; k < NPLANES; k++) dst.at(k, d) = src.at(k, s);
  dst.id[d] = src.id[s];
  dst.density[d] = src.density[s];
}

// ---------------------------------------------------------------------------
// Host launchers.
static inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

template <int SL>
static void mass_tiled(const Params &P, const PState &S, const GState &g, int32_t *boxes,
                       uint32_t wmax, const KeyOut &ko, hipStream_t st) {
  uint32_t nb = nblk(P.n, 64u * SL);
  size_t shm = (size_t)wmax * 4;
  if (P.unit_mass) k_p2g_mass_tiled<SL, true><<<nb, kTB, shm, st>>>(P, S, g, boxes, wmax, ko);
  else k_p2g_mass_tiled<SL, false><<<nb, kTB, shm, st>>>(P, S, g, boxes, wmax, ko);
}
template <int SL>
static void mom_tiled(const Params &P, const PState &S, const GState &g, const int32_t *boxes,
                      uint32_t wmax, hipStream_t st) {
  uint32_t nb = nblk(P.n, 64u * SL);
  size_t shm = (size_t)wmax * 16;
  if (P.unit_mass) k_p2g_mom_tiled<SL, true><<<nb, kTB, shm, st>>>(P, S, g, boxes, wmax);
  else k_p2g_mom_tiled<SL, false><<<nb, kTB, shm, st>>>(P, S, g, boxes, wmax);
}

bool tiled_slices_supported(uint32_t sl) { return sl == 4 || sl == 8 || sl == 16 || sl == 32; }

void launch_p2g_mass_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                           int32_t *boxes, const KeyOut &ko, hipStream_t st) {
  if (!P.n) return;
  switch (tc.slices) {
    case 4: mass_tiled<4>(P, S, g, boxes, tc.wmax, ko, st); break;
    case 8: mass_tiled<8>(P, S, g, boxes, tc.wmax, ko, st); break;
    case 16: mass_tiled<16>(P, S, g, boxes, tc.wmax, ko, st); break;
    default: mass_tiled<32>(P, S, g, boxes, tc.wmax, ko, st); break;
  }
}

void launch_p2g_mom_tiled(const Params &P, const PState &S, const GState &g, const TileCfg &tc,
                          const int32_t *boxes, hipStream_t st) {
  if (!P.n) return;
  switch (tc.slices) {
    case 4: mom_tiled<4>(P, S, g, boxes, tc.wmax, st); break;
    case 8: mom_tiled<8>(P, S, g, boxes, tc.wmax, st); break;
    case 16: mom_tiled<16>(P, S, g, boxes, tc.wmax, st); break;
    default: mom_tiled<32>(P, S, g, boxes, tc.wmax, st); break;
  }
}

KeyOut key_out(const SortBufs &sb, uint32_t tdx, uint32_t tdy) {
  KeyOut ko;
  ko.counts = sb.counts;
  ko.keys = sb.keys;
  ko.rank = sb.rank;
  ko.tile_rank = sb.tile_rank;
  ko.tdx = tdx;
  ko.tdy = tdy;
  return ko;
}

// Full sort (data moved by a scatter pass); used for the first step after a
// reset, when the particles are still in caller order.  The histogram array
// is left zeroed.
void launch_sort(const Params &P, const PState &src, const PState &dst, const SortBufs &sb,
                 uint32_t tdx, uint32_t tdy, uint32_t slices, hipStream_t st) {
  if (!P.n) return;
  k_sort_keys<<<nblk(P.n, 256), 256, 0, st>>>(P, src, key_out(sb, tdx, tdy));
  launch_scan(sb.counts, sb.nkeys, false, sb.offsets, sb.scratch, nullptr, st, sb.counts);
  k_sort_scatter<<<nblk(P.n, 256), 256, 0, st>>>(P.n, slices, src, dst, sb.keys, sb.rank,
                                                 sb.offsets);
}

// After a P2G mass pass that recorded keys: offsets and the inverse
// permutation for k_g2p_relayout.  Leaves the histogram zeroed.
void launch_relayout_prep(const Params &P, const SortBufs &sb, uint32_t slices, hipStream_t st) {
  if (!P.n) return;
  launch_scan(sb.counts, sb.nkeys, false, sb.offsets, sb.scratch, nullptr, st, sb.counts);
  k_build_src<<<nblk(P.n, 256), 256, 0, st>>>(P.n, slices, sb.keys, sb.rank, sb.offsets,
                                              sb.src_of);
}

}  // namespace mpm
