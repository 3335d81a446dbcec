// Blocked-CSR SpMV building blocks shared by smh_bsrmv and the CG kernels.
//
// Reference loop (src/lower/index_expressions/lower_indexexprs.cpp:630-667):
//   for i: tmp = 0; for ij in [rowptr[i], rowptr[i+1]): j = colidx[ij];
//          for bi, bj: tmp[bi] += A[ij*b*b + bi*b + bj] * x[j*b + bj]
// The reference GPU path ran one thread per row over that serial loop, with
// 16-thread blocks (gpu_backend.h:49-51).
//
// MI355X layout of the work ("row tiles"): the stored blocks are cut into
// tiles of whole rows holding at most NB blocks: tile k holds the rows
// whose first block lies in [k*T', (k+1)*T'), T' = NB - maxrow + 1.  One
// persistent workgroup streams a contiguous range of tiles:
//   1. lane (s, bi) owns block row bi of blocks s, s+S, s+2S, s+3S of a tile
//      (S = blocks per round): it reads those b contiguous values each, so a
//      wavefront's loads cover one contiguous range of the value array;
//   2. it gathers x[colidx*b .. +b] (L2/MALL resident) and writes one partial
//      row sum per block into LDS;
//   3. the next tile's values and columns are issued into registers right
//      away, so HBM reads stay in flight while this tile finishes;
//   4. one thread per (row, component) adds its blocks' partial sums in
//      ascending order, then one thread per row hands the row to an epilogue
//      (store, dot partial, residual, ...).
// Tiles are dealt so that each XCD (workgroups go round-robin over 8 XCDs)
// owns a contiguous eighth of the rows and its L2 holds the x entries those
// rows gather.
#pragma once
#include <type_traits>

#include "smh_internal.h"

namespace smh {

// Tile shapes (measured on MI355X at 96^3 x 5 tets: one wavefront of 21
// 3x3 blocks per round and 16 rounds per tile beat three wavefronts x 4
// rounds by 5-7%); scalar (1x1) tiles keep 4 rounds of 256 lanes.
#ifndef SMH_TILE_ROUNDS3
#define SMH_TILE_ROUNDS3 16
#endif
#ifndef SMH_TILE_VEC1
#define SMH_TILE_VEC1 4
#endif
#ifndef SMH_TILE_ROUNDS1
#define SMH_TILE_ROUNDS1 1
#endif
// Rounds per lane of the FP8 3x3 correction pass (one block per lane),
// measured at 96^3 x 5 tets with the row sums on one thread per row: 54 / 49
// / 47 / 60 / 54 / 59 us at 4 / 6 / 8 / 10 / 12 / 16 rounds (6 and 10 gather
// in groups of three and two instead of four; with (row, entry) sums staged
// through LDS, 4 rounds had been best: 58 us)
#ifndef SMH_TILE_ROUNDS8
#define SMH_TILE_ROUNDS8 8
#endif
// matrix values are read once per product: streaming (non-temporal) loads
#ifndef SMH_STREAM_LOAD
#define SMH_STREAM_LOAD(p) __builtin_nontemporal_load(p)
#endif

#ifndef SMH_TILE_THREADS3
#define SMH_TILE_THREADS3 64
#endif
// A tile shape: ROUNDS rounds per lane, THREADS threads per workgroup.
template <int BD, int ROUNDS, int THREADS>
struct TileShapeR {
  static constexpr int kThreads = THREADS;
  static constexpr int kRounds = ROUNDS;                   // per lane
  static constexpr int kWaves = kThreads / 64;
  static constexpr int kVec = BD == 1 ? SMH_TILE_VEC1 : 1;  // consecutive blocks per lane
  static constexpr int kPerWave = 64 / BD * kVec;         // blocks per wave and round
  static constexpr int kPerRound = kWaves * kPerWave;     // blocks per round
  static constexpr int kBlocks = kPerRound * kRounds;     // blocks per tile
  static constexpr bool kBPL = false;                     // one block per lane
  static constexpr bool kPlanar = false;                  // lane-planar FP32 values
  static_assert(kThreads % 64 == 0, "tiles are made of whole wavefronts");
};
// the default shape of a block size
template <int BD>
struct TileShape : TileShapeR<BD, BD == 3 ? SMH_TILE_ROUNDS3 : SMH_TILE_ROUNDS1,
                              BD == 3 ? SMH_TILE_THREADS3 : 256> {};
// FP8 3x3 blocks (the learned correction) put a whole block on one lane: its
// three 4-byte row words are one 12-byte load and 64 blocks go per wave and
// round, instead of one word per lane and 21 blocks.
struct TileShapeB {
  static constexpr int kThreads = SMH_TILE_THREADS3;
  static constexpr int kRounds = SMH_TILE_ROUNDS8;
  static constexpr int kWaves = kThreads / 64;
  static constexpr int kVec = 1;
  static constexpr int kPerWave = 64;
  static constexpr int kPerRound = kWaves * kPerWave;
  static constexpr int kBlocks = kPerRound * kRounds;
  static constexpr bool kBPL = true;
  static constexpr bool kPlanar = false;
};

// Lane-planar FP32 3x3 values (SMH_F32P): the blocks are cut into groups of
// 64 consecutive blocks, and group g occupies the 576 floats at 576*g:
//   [64 x float4: entries 0..3 of each block][64 x float4: entries 4..7]
//   [64 x float: entry 8]
// (entry e = 3*bi + bj of the block, as in the row-major layout).  One lane
// takes one block per round: two 16-byte loads and one 4-byte load, each of
// them one contiguous run of the wave's 64 blocks (at most two runs when the
// tile starts inside a group), against one 12-byte load per block row in the
// row-major layout.  Measured at 96^3 x 5 tets (FP32, 11.7M blocks): 91 us
// against 104 us for the same products, bitwise equal (each block row is
// summed in the same order).  Rounds per lane (tiles of 64 x rounds
// blocks), measured at 96^3 x 5 tets: the plain FP32 product 88.1-88.8 us at
// 8 rounds across boxes (87.6-92.3 at 12, 93.9 at 10, 99.1 at 16); the
// solvers' SpMV with FP64 row sums (TileShapePlS) 96-98 us at 12 rounds
// against 104-106 at 8, 107 at 10 and 129 at 16 (12 rounds hold 220
// registers, 2 waves per SIMD).
#ifndef SMH_TILE_ROUNDS_PL
#define SMH_TILE_ROUNDS_PL 8
#endif
#ifndef SMH_TILE_ROUNDS_PLS
#define SMH_TILE_ROUNDS_PLS 12
#endif
// Grid of the planar kernels.  Their workgroups stride over their XCD's
// eighth of the tiles (TileDev::interleave), so the tiles in flight on one
// XCD form one moving window (its L2 holds the x they gather) and the load
// stays even for any grid.  Measured at 96^3 x 5 tets: SpMV 91 -> 87 us at
// 8,192 workgroups, 86.4-87.4 us at 16,384 against 89.1-89.6 at 8,192 on one
// box (the solvers' copy has its own grid, SMH_TILE_GRID_PLS: their update
// kernel sums one partial per SpMV workgroup).
#ifndef SMH_TILE_GRID_PL
#define SMH_TILE_GRID_PL 16384
#endif
constexpr uint64_t kPlanarGroup = 576;  // floats per group of 64 blocks
__host__ __device__ inline uint64_t planar_floats(uint64_t nnz) {
  return ((nnz + 63) / 64) * kPlanarGroup;
}
template <int ROUNDS>
struct TileShapePlR {
  static constexpr int kThreads = 64;
  static constexpr int kRounds = ROUNDS;
  static constexpr int kWaves = 1;
  static constexpr int kVec = 1;
  static constexpr int kPerWave = 64;
  static constexpr int kPerRound = 64;
  static constexpr int kBlocks = kPerRound * kRounds;
  static constexpr bool kBPL = true;
  static constexpr bool kPlanar = true;
};
using TileShapePl = TileShapePlR<SMH_TILE_ROUNDS_PL>;    // smh_bsrmv (FP32 row sums)
using TileShapePlS = TileShapePlR<SMH_TILE_ROUNDS_PLS>;  // the solvers' inner SpMV
inline uint32_t tile_blocks_host(int bd) {
  return bd == 3 ? (SMH_TILE_THREADS3 / 64) * 21 * SMH_TILE_ROUNDS3
                 : 256 * SMH_TILE_VEC1 * SMH_TILE_ROUNDS1;
}

// column index of a lane that gathers nothing in a round
constexpr uint32_t kNoCol = 0xFFFFFFFFu;

// N (4) consecutive entries at p as 16-byte loads (dword-aligned addresses are
// enough for global loads); other types element by element
template <typename S, int N, bool NT>
__device__ __forceinline__ void load_vec(const S *p, S *out) {
  if constexpr (N == 4 && sizeof(S) == 4) {
    typedef S v4 __attribute__((ext_vector_type(4), aligned(4)));
    v4 r;
    if constexpr (NT) r = __builtin_nontemporal_load((const v4 *)p);
    else r = *(const v4 *)p;
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
  } else if constexpr (N == 4 && std::is_same<S, double>::value) {
    typedef double v2 __attribute__((ext_vector_type(2), aligned(8)));
    v2 r0, r1;
    if constexpr (NT) {
      r0 = __builtin_nontemporal_load((const v2 *)p);
      r1 = __builtin_nontemporal_load((const v2 *)(p + 2));
    } else {
      r0 = *(const v2 *)p;
      r1 = *(const v2 *)(p + 2);
    }
    out[0] = r0.x;
    out[1] = r0.y;
    out[2] = r1.x;
    out[3] = r1.y;
  } else {
#pragma unroll
    for (int b = 0; b < N; ++b) {
      if constexpr (NT) out[b] = __builtin_nontemporal_load(p + b);
      else out[b] = p[b];
    }
  }
}

// BD consecutive entries at p in as few loads as fit them exactly; NT:
// non-temporal (streamed once).  Three floats: one 12-byte load (a 3-vector
// type with its alignment lowered to 4 reads exactly 12 bytes).  Three
// doubles: 16 + 8 bytes (a 3-vector of doubles is lowered to two 16-byte
// loads, which would read past the row).
template <typename S, int BD, bool NT>
__device__ __forceinline__ void load_row(const S *p, S *out) {
  if constexpr (BD == 3 && std::is_same<S, float>::value) {
    typedef float v3 __attribute__((ext_vector_type(3), aligned(4)));
    v3 r;
    if constexpr (NT) r = __builtin_nontemporal_load((const v3 *)p);
    else r = *(const v3 *)p;
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
  } else if constexpr (BD == 3 && std::is_same<S, double>::value) {
    typedef double v2 __attribute__((ext_vector_type(2), aligned(8)));
    v2 r;
    if constexpr (NT) {
      r = __builtin_nontemporal_load((const v2 *)p);
      out[2] = __builtin_nontemporal_load(p + 2);
    } else {
      r = *(const v2 *)p;
      out[2] = p[2];
    }
    out[0] = r.x;
    out[1] = r.y;
  } else {
#pragma unroll
    for (int b = 0; b < BD; ++b) {
      if constexpr (NT) out[b] = __builtin_nontemporal_load(p + b);
      else out[b] = p[b];
    }
  }
}
inline uint32_t tile_threads_host(int bd) { return bd == 3 ? SMH_TILE_THREADS3 : 256; }

template <int PREC> struct Val;
template <> struct Val<SMH_F64> {
  using T = double;  // vector / compute type
  using S = double;  // stored value type
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = false;
};
template <> struct Val<SMH_F32> {
  using T = float;
  using S = float;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = false;
};
template <> struct Val<SMH_F32X2> {
  using T = double;
  using S = float;
  static constexpr bool kLo = true;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = false;
};
constexpr int kF32D = 3;  // float values driven by double vectors
template <> struct Val<kF32D> {
  using T = double;
  using S = float;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = false;
};
constexpr int kBF16F = 4;  // bfloat16 values (bits in uint16), float vectors
template <> struct Val<kBF16F> {
  using T = float;
  using S = uint16_t;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = true;
  static constexpr bool kFP8 = false;
};
constexpr int kBF16D = 5;  // bfloat16 values, double vectors
template <> struct Val<kBF16D> {
  using T = double;
  using S = uint16_t;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = true;
  static constexpr bool kFP8 = false;
};

// 8-bit floats (e4m3: bias 7, 3 mantissa bits, subnormals below 2^-6) of
// 3x3 blocks stored as 3 block rows of 4 bytes (3 values and a zero pad), so
// one lane reads one aligned 32-bit word per block row.  The stored values
// are the block's entries times 2^5 (kFP8Scale undoes that).
constexpr int kFP8F = 6;  // FP8 values, float vectors
constexpr int kFP8D = 7;  // FP8 values, double vectors
constexpr float kFP8Scale = 1.0f / 32.0f;
template <> struct Val<kFP8F> {
  using T = float;
  using S = uint32_t;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = true;
};
template <> struct Val<kFP8D> {
  using T = double;
  using S = uint32_t;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = true;
};
__host__ __device__ __forceinline__ float e4m3_to_float(uint32_t b) {
  const uint32_t e = (b >> 3) & 15u, m = b & 7u;
#ifdef __HIP_DEVICE_COMPILE__
  const float v = e ? __uint_as_float(((e + 120u) << 23) | (m << 20)) : (float)m * 0.001953125f;
#else
  float v;
  if (e) {
    const uint32_t bits = ((e + 120u) << 23) | (m << 20);
    __builtin_memcpy(&v, &bits, 4);
  } else {
    v = (float)m * 0.001953125f;
  }
#endif
  return (b & 0x80u) ? -v : v;
}

// FP8 e4m3 byte (low 8 bits of b) -> float times kFP8Scale, without branches:
// the sign moved to bit 31 and the seven magnitude bits to bits 26..20 read
// as 2^(e-127) (1 + m/8) for e > 0 and as the FP32 subnormal m 2^-129 for
// e = 0 (kernels run with FP32 subnormals preserved); one exact multiply by
// 2^120 * kFP8Scale gives the e4m3 value times kFP8Scale.  (A select between
// the normal and subnormal forms compiles to branches that serialise the
// tile loop's gathers.)
__device__ __forceinline__ float e4m3_scaled(uint32_t b) {
  return __uint_as_float(((b & 0x80u) << 24) | ((b & 0x7Fu) << 20)) * (0x1p120f * kFP8Scale);
}

// IEEE half values (bits in uint16): the learned correction in stencil
// storage, read by one thread per row as 2-byte loads
constexpr int kF16F = 8;  // FP16 values, float vectors
constexpr int kF16D = 9;  // FP16 values, double vectors
template <> struct Val<kF16F> {
  using T = float;
  using S = uint16_t;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = false;
};
template <> struct Val<kF16D> {
  using T = double;
  using S = uint16_t;
  static constexpr bool kLo = false;
  static constexpr bool kBF16 = false;
  static constexpr bool kFP8 = false;
};

// stored value -> compute type
template <int PREC>
__device__ __forceinline__ typename Val<PREC>::T val_of(typename Val<PREC>::S v) {
  if constexpr (PREC == kF16F || PREC == kF16D)
    return (typename Val<PREC>::T)(float)__builtin_bit_cast(_Float16, (uint16_t)v);
  else if constexpr (Val<PREC>::kBF16)
    return (typename Val<PREC>::T)__uint_as_float((uint32_t)v << 16);
  else
    return (typename Val<PREC>::T)v;
}

// the tile shape a precision runs with
template <int PREC, int BD>
using TileShapeP =
    typename std::conditional<Val<PREC>::kFP8 && BD == 3, TileShapeB, TileShape<BD>>::type;
template <int PREC, int BD>
inline uint32_t tile_blocks_of() {
  return (uint32_t)TileShapeP<PREC, BD>::kBlocks;
}

struct TileDev {
  uint32_t nrows;
  uint64_t nnz;
  const uint32_t *rowptr;
  const uint32_t *colidx;
  const void *vals;
  const void *vals_lo;
  const uint32_t *tile_row;  // [ntiles+1]
  uint32_t ntiles;
  uint32_t tiles_per_wg;
  uint32_t interleave;  // 1: workgroups of an XCD stride over its tiles (see wg_tiles)
  // batched systems (optional): the system of each row and the systems'
  // states; a tile whose rows all belong to systems that are no longer
  // iterating (state != 0) is streamed as an empty tile
  const uint32_t *row_sys = nullptr;
  const int *sys_state = nullptr;
};

template <int PREC, int BD, class SH = TileShapeP<PREC, BD>>
struct TileSmem {
  using T = typename Val<PREC>::T;
  static constexpr int kBlocks = SH::kBlocks;
  // FP32 sums over lane-planar tiles stage each (row, entry) sum in rsum
  // before the epilogue; the other loops sum a whole row on one thread
  static constexpr bool kStaged = sizeof(T) == 4 && SH::kPlanar;
  T part[kBlocks * BD];
  T rsum[kStaged ? (kBlocks + 1) * BD : 1];
  uint32_t rp[kBlocks + 2];  // rowptr of the tile's rows, relative
  uint32_t bnd[2][4];        // R0, R1, base, end of this and the next tile
  double red[SH::kWaves];
};

// Gather of x for the tile loop: x_j as BD consecutive values.
template <typename T>
struct XPlain {
  const T *x;
  int bd;
  __device__ T operator()(uint32_t j, int b) const { return x[(uint64_t)j * bd + b]; }
  // x_j (BD entries) as one vector load
  template <int BD>
  __device__ void row(uint32_t j, T *out) const {
    load_row<T, BD, false>(x + (uint64_t)j * BD, out);
  }
};

// Tile kernels run persistent workgroups (16 per CU on 256 CUs by default),
// each over a contiguous range of tiles; at most kMaxTileGrid of them.
constexpr int kMaxTileGrid = 16384;
// Grid of the learned correction pass (8,192: 53.4 us against 54.3 at 4
// rounds per tile, 48.0 against 47.0 at 8)
#ifndef SMH_TILE_GRID_C
#define SMH_TILE_GRID_C 4096
#endif
// Grid of the CG x/r update, whose every workgroup sums the SpMV's partials:
// 512 / 1,024 / 2,048 workgroups took 22.6 / 23.4 / 24.2 us at 96^3 x 5 tets
// and 91.5 / 74.2 / 82.5 us on the 192^3 hex system
#ifndef SMH_CG_UPDATE_GRID
#define SMH_CG_UPDATE_GRID 1024
#endif
// Grid of the CG p update and the other vector passes: FP32 vectors 2,048
// (1,024: 6.7 against 6.0 us at 96^3 x 5 tets), FP64 vectors 1,024 (27.6
// against 30.5 us on the 192^3 hex system)
#ifndef SMH_CG_VEC_GRID
#define SMH_CG_VEC_GRID 2048
#endif
#ifndef SMH_CG_VEC_GRID64
#define SMH_CG_VEC_GRID64 1024
#endif
// Grid of the solvers' planar inner SpMV at 96^3 x 5 tets: 8,192 / 12,288 /
// 16,384 workgroups 98.4 / 95.0 / 95.4 us, the update kernel that sums their
// partials 23.4 / 24.6 / 28.3 us
#ifndef SMH_TILE_GRID_PLS
#define SMH_TILE_GRID_PLS 12288
#endif
// Grid of the stencil row kernels on the 192^3 hex system: 2,048 / 4,096 /
// 8,192 workgroups, the solver's SpMV 206.7 / 204.4 / 198.3 us, its
// correction pass 121.8 / 120.6 / 117.2 us
#ifndef SMH_ST_GRID
#define SMH_ST_GRID 8192
#endif
#ifndef SMH_TILE_GRID
#define SMH_TILE_GRID 4096
#endif
static_assert(SMH_TILE_GRID <= kMaxTileGrid, "tile grid above kMaxTileGrid");
static_assert(SMH_TILE_GRID_PLS <= kMaxTileGrid && SMH_TILE_GRID_C <= kMaxTileGrid &&
                  SMH_ST_GRID <= kMaxTileGrid,
              "solver grids above kMaxTileGrid");
static_assert(SMH_TILE_GRID_PL <= kMaxTileGrid, "tile grid above kMaxTileGrid");

// Scalar (1x1) products of a single system stride over the tiles (below):
// the 27-point rows of a structured 3D mesh reach +-n^2 rows away, and
// contiguous per-workgroup ranges spread an XCD's active rows over its whole
// chunk, whose x no longer fits that XCD's L2.  Measured on the 192^3 hex
// system: SpMV 485 -> 481 us, a Jacobi-PCG iteration 665 -> 654 us (the
// re-reads mostly hit the Infinity Cache); results are bitwise the same.
#ifndef SMH_TILE_INTERLEAVE1
#define SMH_TILE_INTERLEAVE1 1
#endif

inline TileDev tile_dev(const smh_bsr *A, const smh_tile_plan *p, unsigned *grid,
                        bool interleave = false, unsigned cap = SMH_TILE_GRID) {
  uint64_t g = p->ntiles < (uint64_t)cap ? p->ntiles : (uint64_t)cap;
  if (g >= 8) g &= ~7ull;
  if (g < 1) g = 1;
  *grid = (unsigned)g;
  TileDev d{A->nrows, A->nnz, A->rowptr, A->colidx, A->vals, A->vals_lo, p->tile_row, p->ntiles,
            (uint32_t)((p->ntiles + g - 1) / g), (interleave && g >= 8) ? 1u : 0u,
            nullptr, nullptr};
  return d;
}

// Map workgroup g to its tile range (XCD-contiguous when the grid is a
// multiple of 8).  Interleaved: XCD x (workgroups g = x mod 8) owns the x-th
// eighth of the tiles and its m = grid/8 workgroups take every m-th tile of
// it, so the tiles in flight on one XCD form one moving window; the tile
// loop then steps by m (tile_step).
__device__ __forceinline__ uint32_t tile_step(const TileDev &A) {
  return A.interleave ? gridDim.x >> 3 : 1u;
}
__device__ __forceinline__ void wg_tiles(const TileDev &A, uint32_t &k0, uint32_t &k1) {
  const uint32_t nb = gridDim.x;
  if (A.interleave) {
    const uint32_t chunk = (A.ntiles + 7) >> 3;
    const uint64_t lo = (uint64_t)(blockIdx.x & 7) * chunk;
    const uint64_t hi = min<uint64_t>(lo + chunk, A.ntiles);
    k0 = (uint32_t)min<uint64_t>(lo + (blockIdx.x >> 3), hi);
    k1 = (uint32_t)hi;
    return;
  }
  uint32_t lb = blockIdx.x;
  if ((nb & 7) == 0) lb = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
  const uint64_t a = (uint64_t)lb * A.tiles_per_wg;
  k0 = (uint32_t)min<uint64_t>(a, A.ntiles);
  k1 = (uint32_t)min<uint64_t>(a + A.tiles_per_wg, A.ntiles);
}

// Registers holding one tile's values and columns for one lane.
//
// Lane layout within a wavefront, per round: block s of the wave's kPerWave
// blocks is held by lanes BD*s .. BD*s+BD-1, one block row each (values read
// fully coalesced).  For BD > 1 the column index and the gathered x of block
// s are loaded once, by lane s, and handed to the block's lanes by a
// cross-lane shuffle: the contiguous gathering lanes keep the address unit's
// work per block at one column and one x request instead of BD of each.
template <int PREC, int BD, class SHP = TileShape<BD>>
struct TileRegs {
  using S = typename Val<PREC>::S;
  static constexpr int kRounds = SHP::kRounds;
  static constexpr int kV = Val<PREC>::kFP8 ? 1 : BD;  // FP8: one word per block row
  S v[kRounds][kV];
  float vl[Val<PREC>::kLo ? kRounds : 1][BD];
  uint32_t col[kRounds];  // the column this lane gathers (kNoCol: none)

  __device__ __forceinline__ void load(const TileDev &A, uint32_t base, uint32_t end, int w,
                                       int lane) {
    using SH = SHP;
    constexpr int BB = BD * BD;
    const int s = w * SH::kPerWave + lane / BD, bi = lane % BD;
    const bool ok = lane < SH::kPerWave * BD;
    if (base >= end) {  // a tile without blocks (the matrix may have none)
#pragma unroll
      for (int u = 0; u < kRounds; ++u) {
        col[u] = kNoCol;
#pragma unroll
        for (int b = 0; b < kV; ++b) v[u][b] = (S)0;
        if constexpr (Val<PREC>::kLo)
#pragma unroll
          for (int b = 0; b < BD; ++b) vl[u][b] = 0.0f;
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < kRounds; ++u) {
      const uint32_t blk = base + s + u * SH::kPerRound;
      const bool act = ok && blk < end;
      const uint32_t gblk = base + w * SH::kPerWave + lane + u * SH::kPerRound;
      col[u] = (lane < SH::kPerWave && gblk < end) ? A.colidx[gblk] : kNoCol;
      // 4-byte values: lanes without a block skip the load (measured ~7%
      // faster on FP32 than loading a dummy row).  8-byte values: branch-free
      // (below); the conditional form makes the compiler keep FP64 rows in
      // scratch memory.
      if constexpr (sizeof(S) < 8) {
        if constexpr (Val<PREC>::kFP8) {
          v[u][0] = act ? SMH_STREAM_LOAD((const uint32_t *)A.vals + (uint64_t)blk * BD + bi) : 0u;
          continue;
        }
        const uint64_t o = (uint64_t)blk * BB + bi * BD;
        if (act) {
          load_row<S, BD, true>((const S *)A.vals + o, v[u]);
          if constexpr (Val<PREC>::kLo) load_row<float, BD, true>((const float *)A.vals_lo + o, vl[u]);
        } else {
#pragma unroll
          for (int b = 0; b < kV; ++b) v[u][b] = (S)0;
          if constexpr (Val<PREC>::kLo)
#pragma unroll
            for (int b = 0; b < BD; ++b) vl[u][b] = 0.0f;
        }
        continue;
      }
      // branch-free: a lane without a block reads row 0 of block 0 (present
      // whenever the tile has blocks) and keeps zeros
      const uint64_t o = act ? (uint64_t)blk * BB + bi * BD : 0;
      S t[BD];
      load_row<S, BD, true>((const S *)A.vals + o, t);
#pragma unroll
      for (int b = 0; b < BD; ++b) v[u][b] = act ? t[b] : (S)0;
      if constexpr (Val<PREC>::kLo) {
        float tl[BD];
        load_row<float, BD, true>((const float *)A.vals_lo + o, tl);
#pragma unroll
        for (int b = 0; b < BD; ++b) vl[u][b] = act ? tl[b] : 0.0f;
      }
    }
  }
};

// Registers of one lane for 1x1 tiles: kVec consecutive entries per round,
// values and columns loaded as vectors (a 4-byte value per request would
// leave the address unit, not HBM, as the limit).
template <int PREC, class SHP = TileShape<1>>
struct TileRegsV {
  using S = typename Val<PREC>::S;
  using SH = SHP;
  static constexpr int kRounds = SH::kRounds, kVec = SH::kVec;
  S v[kRounds][kVec];
  float vl[Val<PREC>::kLo ? kRounds : 1][kVec];
  uint32_t col[kRounds][kVec];  // kNoCol: past the tile's end

  __device__ __forceinline__ void load(const TileDev &A, uint32_t base, uint32_t end, int w,
                                       int lane) {
#pragma unroll
    for (int u = 0; u < kRounds; ++u) {
      const uint32_t b0 = base + u * SH::kPerRound + (uint32_t)(w * 64 + lane) * kVec;
      if (b0 + kVec <= end) {
        load_vec<S, kVec, true>((const S *)A.vals + b0, v[u]);
        load_vec<uint32_t, kVec, false>(A.colidx + b0, col[u]);
        if constexpr (Val<PREC>::kLo)
          load_vec<float, kVec, true>((const float *)A.vals_lo + b0, vl[u]);
      } else {
#pragma unroll
        for (int j = 0; j < kVec; ++j) {
          const bool a = b0 + j < end;
          v[u][j] = a ? ((const S *)A.vals)[b0 + j] : (S)0;
          col[u][j] = a ? A.colidx[b0 + j] : kNoCol;
          if constexpr (Val<PREC>::kLo) vl[u][j] = a ? ((const float *)A.vals_lo)[b0 + j] : 0.0f;
        }
      }
    }
  }
};

// Registers of one lane for FP8 3x3 tiles (TileShapeB): one whole block per
// round, its three row words and its column.  (The same layout for FP32
// blocks, nine values per lane, measured slower than one block row per lane:
// 126 us against 106 us at 96^3 x 5.)
template <int PREC>
struct TileRegsB {
  using SH = TileShapeB;
  static_assert(Val<PREC>::kFP8, "one block per lane: FP8 blocks");
  static constexpr int kRounds = SH::kRounds;
  uint32_t v[kRounds][3];
  uint32_t col[kRounds];  // kNoCol: past the tile's end

  __device__ __forceinline__ void load(const TileDev &A, uint32_t base, uint32_t end, int w,
                                       int lane) {
#pragma unroll
    for (int u = 0; u < kRounds; ++u) {
      const uint32_t blk = base + u * SH::kPerRound + (uint32_t)(w * 64 + lane);
      if (blk < end) {
        load_row<uint32_t, 3, true>((const uint32_t *)A.vals + (uint64_t)blk * 3, v[u]);
        col[u] = A.colidx[blk];
      } else {
        v[u][0] = v[u][1] = v[u][2] = 0u;
        col[u] = kNoCol;
      }
    }
  }
};

// Registers of one lane for lane-planar FP32 3x3 tiles (TileShapePl,
// TileShapePlS): one whole block per round, its nine values and its column.
template <int PREC, class SH>
struct TileRegsPl {
  static_assert(std::is_same<typename Val<PREC>::S, float>::value && !Val<PREC>::kLo,
                "lane-planar tiles hold plain float values");
  static constexpr int kRounds = SH::kRounds;
  float v[kRounds][9];
  uint32_t col[kRounds];  // kNoCol: past the tile's end

  __device__ __forceinline__ void load(const TileDev &A, uint32_t base, uint32_t end, int w,
                                       int lane) {
    typedef float v4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int u = 0; u < kRounds; ++u) {
      const uint32_t blk = base + u * SH::kPerRound + (uint32_t)(w * 64 + lane);
      if (blk < end) {
        const float *g = (const float *)A.vals + (uint64_t)(blk >> 6) * kPlanarGroup;
        const uint32_t l = blk & 63u;
        const v4 a = __builtin_nontemporal_load((const v4 *)(g + 4 * l));
        const v4 b = __builtin_nontemporal_load((const v4 *)(g + 256 + 4 * l));
        v[u][0] = a.x;
        v[u][1] = a.y;
        v[u][2] = a.z;
        v[u][3] = a.w;
        v[u][4] = b.x;
        v[u][5] = b.y;
        v[u][6] = b.z;
        v[u][7] = b.w;
        v[u][8] = __builtin_nontemporal_load(g + 512 + l);
        col[u] = A.colidx[blk];
      } else {
#pragma unroll
        for (int e = 0; e < 9; ++e) v[u][e] = 0.0f;
        col[u] = kNoCol;
      }
    }
  }
};

// Tile boundaries (R0, R1, base, end) of tile k into slot.
__device__ __forceinline__ void tile_bounds(const TileDev &A, uint32_t k, uint32_t *slot) {
  const uint32_t R0 = A.tile_row[k], R1 = A.tile_row[k + 1];
  slot[0] = R0;
  slot[1] = R1;
  slot[2] = A.rowptr[R0];
  slot[3] = A.rowptr[R1];
  if (A.row_sys && R1 > R0) {
    bool active = false;
    for (uint32_t m = A.row_sys[R0]; m <= A.row_sys[R1 - 1] && !active; ++m)
      active = A.sys_state[m] == 0;
    if (!active) {  // no rows, no blocks: the tile costs its loop turn only
      slot[1] = slot[0];
      slot[3] = slot[2];
    }
  }
}

// Stream tiles [k0, k1): rows handed to epi(row, s[BD]) by one thread per
// row.  XG gathers x: xg(j, bj).
template <int PREC, int BD, class SH, class XG, class EPI>
__device__ __forceinline__ void spmv_tiles(const TileDev &A, uint32_t k0, uint32_t k1,
                                           TileSmem<PREC, BD, SH> &sm, const XG &xg, EPI &epi) {
  using T = typename Val<PREC>::T;
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int s = SH::kBPL ? w * 64 + lane : w * SH::kPerWave + lane / BD, bi = lane % BD;
  const bool ok = SH::kBPL || lane < SH::kPerWave * BD;
  if (k0 >= k1) return;
  const uint32_t ks = tile_step(A);
  if (t == 0) {
    tile_bounds(A, k0, sm.bnd[0]);
    if (k0 + ks < k1) tile_bounds(A, k0 + ks, sm.bnd[1]);
  }
  __syncthreads();
  using Regs = typename std::conditional<
      BD == 1, TileRegsV<PREC, SH>,
      typename std::conditional<
          SH::kPlanar, TileRegsPl<PREC, SH>,
          typename std::conditional<SH::kBPL, TileRegsB<PREC>,
                                    TileRegs<PREC, BD, SH>>::type>::type>::type;
  Regs cur, nxt;
  cur.load(A, sm.bnd[0][2], sm.bnd[0][3], w, lane);
  for (uint32_t k = k0; k < k1; k += ks) {
    const uint32_t R0 = sm.bnd[0][0], R1 = sm.bnd[0][1], base = sm.bnd[0][2];
    const uint32_t nbase = sm.bnd[1][2], nend = sm.bnd[1][3];
    const uint32_t nr = R1 - R0;
    const bool small = nr + 1 <= (uint32_t)(SH::kBlocks + 1);
    // read ahead the boundaries of tile k+2
    uint32_t ahead[4] = {0, 0, 0, 0};
    if (t == 0 && k + 2 * ks < k1) tile_bounds(A, k + 2 * ks, ahead);
    if (small)
      for (uint32_t i = t; i <= nr; i += SH::kThreads) sm.rp[i] = A.rowptr[R0 + i] - base;
    // partial row sums of this lane's blocks
    if constexpr (BD == 1) {
#pragma unroll
      for (int u = 0; u < SH::kRounds; ++u) {
        constexpr int kVec = SH::kVec;
        T xo[kVec];
#pragma unroll
        for (int q = 0; q < kVec; ++q) {
          const uint32_t c = cur.col[u][q];
          xo[q] = xg(c == kNoCol ? 0u : c, 0);
        }
        const uint32_t s0 = u * SH::kPerRound + (uint32_t)(w * 64 + lane) * kVec;
#pragma unroll
        for (int q = 0; q < kVec; ++q) {
          T val = val_of<PREC>(cur.v[u][q]);
          if (Val<PREC>::kLo) val += (T)cur.vl[Val<PREC>::kLo ? u : 0][q];
          sm.part[s0 + q] = val * xo[q];
        }
      }
    } else if constexpr (SH::kBPL) {
      // one block per lane: the lane gathers its own x_j, decodes its three
      // row words and forms the block's three row products (each in the
      // order of the one-row-per-lane form, so the results are identical)
      constexpr int kG = SH::kRounds % 4 == 0 ? 4
                         : SH::kRounds % 3 == 0 ? 3
                         : SH::kRounds % 2 == 0 ? 2
                                                : 1;
#pragma unroll
      for (int u0 = 0; u0 < SH::kRounds; u0 += kG) {
        T xo[kG][3];
#pragma unroll
        for (int j = 0; j < kG; ++j) {
          const uint32_t c = cur.col[u0 + j];
          xg.template row<3>(c == kNoCol ? 0u : c, xo[j]);
        }
#pragma unroll
        for (int j = 0; j < kG; ++j) {
          const int u = u0 + j;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            T acc = 0;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
              T val;
              if constexpr (SH::kPlanar)
                val = (T)cur.v[u][3 * a + b];
              else
                val = (T)e4m3_scaled(cur.v[u][a] >> (8 * b));
              acc += val * xo[j][b];
            }
            sm.part[(s + u * SH::kPerRound) * 3 + a] = acc;
          }
        }
      }
    } else {
      // gathers in groups of kG rounds: the group's x loads go out together
      // (a lane past the tile's end gathers row 0, whose product multiplies a
      // zero value and lands in a partial slot no row reads)
      constexpr int kG = SH::kRounds < 4 ? SH::kRounds : 4;
      static_assert(SH::kRounds % kG == 0, "rounds come in whole gather groups");
#pragma unroll
      for (int u0 = 0; u0 < SH::kRounds; u0 += kG) {
        T xo[kG][BD];
        if (BD == 1 || lane < SH::kPerWave) {
#pragma unroll
          for (int j = 0; j < kG; ++j) {
            const uint32_t c = cur.col[u0 + j];
            xg.template row<BD>(c == kNoCol ? 0u : c, xo[j]);
          }
        } else {
#pragma unroll
          for (int j = 0; j < kG; ++j)
#pragma unroll
            for (int b = 0; b < BD; ++b) xo[j][b] = T(0);
        }
#pragma unroll
        for (int j = 0; j < kG; ++j) {
          const int u = u0 + j;
          T xv[BD];
#pragma unroll
          for (int b = 0; b < BD; ++b) {
            if constexpr (BD == 1)
              xv[b] = xo[j][b];
            else
              xv[b] = __shfl(xo[j][b], lane / BD, 64);
          }
          T acc = 0;
#pragma unroll
          for (int b = 0; b < BD; ++b) {
            T val;
            if constexpr (Val<PREC>::kFP8)
              val = (T)e4m3_scaled((uint32_t)cur.v[u][0] >> (8 * b));
            else
              val = val_of<PREC>(cur.v[u][b]);
            if (Val<PREC>::kLo) val += (T)cur.vl[Val<PREC>::kLo ? u : 0][b];
            acc += val * xv[b];
          }
          if (ok) sm.part[(s + u * SH::kPerRound) * BD + bi] = acc;
        }
      }
    }
    // the next tile's loads go out before this tile is reduced
    if (k + ks < k1) nxt.load(A, nbase, nend, w, lane);
    __syncthreads();
    if (small && TileSmem<PREC, BD, SH>::kStaged) {
      // one thread per (row, entry), then one per row for the epilogue
      // (measured faster for the FP32 planar SpMV: 88.6 us against 91.1 us at
      // 96^3 x 5 tets)
      for (uint32_t i = t; i < nr * BD; i += SH::kThreads) {
        const uint32_t r = i / BD, a = i - (i / BD) * BD;
        const uint32_t q0 = sm.rp[r], q1 = sm.rp[r + 1];
        T acc = 0;
        uint32_t q = q0;
        // four independent LDS reads in flight, added in ascending order
        for (; q + 4 <= q1; q += 4) {
          const T v0 = sm.part[q * BD + a], v1 = sm.part[(q + 1) * BD + a];
          const T v2 = sm.part[(q + 2) * BD + a], v3 = sm.part[(q + 3) * BD + a];
          acc += v0;
          acc += v1;
          acc += v2;
          acc += v3;
        }
        for (; q < q1; ++q) acc += sm.part[q * BD + a];
        sm.rsum[i] = acc;
      }
      __syncthreads();
      for (uint32_t i = t; i < nr; i += SH::kThreads) {
        T r[BD];
        for (int a = 0; a < BD; ++a) r[a] = sm.rsum[i * BD + a];
        epi(R0 + i, r);
      }
    } else if (small) {
      // one thread per row sums its BD entries over the row's blocks, each in
      // ascending block order (four blocks' LDS reads in flight, then their
      // adds in order), and hands them to the epilogue: the same sums as the
      // staged form, no second pass (the correction pass 55.9 -> 54.4 us, the
      // FP64 SpMV 257 -> 240 us at 96^3 x 5 tets), half the LDS for FP64 sums
      for (uint32_t i = t; i < nr; i += SH::kThreads) {
        const uint32_t q0 = sm.rp[i], q1 = sm.rp[i + 1];
        T r[BD];
#pragma unroll
        for (int a = 0; a < BD; ++a) r[a] = 0;
        uint32_t q = q0;
        for (; q + 4 <= q1; q += 4) {
          T v[4][BD];
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int a = 0; a < BD; ++a) v[j][a] = sm.part[(q + j) * BD + a];
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int a = 0; a < BD; ++a) r[a] += v[j][a];
        }
        for (; q < q1; ++q)
#pragma unroll
          for (int a = 0; a < BD; ++a) r[a] += sm.part[q * BD + a];
        epi(R0 + i, r);
      }
    } else {
      for (uint32_t i = t; i < nr; i += SH::kThreads) {
        const uint32_t q0 = A.rowptr[R0 + i] - base, q1 = A.rowptr[R0 + i + 1] - base;
        T r[BD];
        for (int a = 0; a < BD; ++a) r[a] = 0;
        for (uint32_t q = q0; q < q1; ++q)
          for (int a = 0; a < BD; ++a) r[a] += sm.part[q * BD + a];
        epi(R0 + i, r);
      }
    }
    __syncthreads();
    if (t == 0) {
      for (int i = 0; i < 4; ++i) {
        sm.bnd[0][i] = sm.bnd[1][i];
        sm.bnd[1][i] = ahead[i];
      }
    }
    cur = nxt;
    __syncthreads();
  }
}

// workgroup-wide FP64 sum (NW wavefronts), fixed order; valid on every thread
template <int NW>
__device__ __forceinline__ double block_sum(double v, double *s) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = s[0];
  for (int w = 1; w < NW; ++w) t += s[w];
  __syncthreads();
  return t;
}
__device__ __forceinline__ double block_sum256(double v, double *s4) { return block_sum<4>(v, s4); }

}  // namespace smh
