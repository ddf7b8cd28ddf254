// Tile binning: intersect keys, a stable on-device LSD radix sort, bin edges,
// and the depth-first binning pipeline used by the fused render path.
//
// Reference semantics: forward.cu:13-98, helpers.cuh:37-51,75-92,
// utils.py:40-59,106-162 (torch.sort is replaced by the stable radix sort).
#include <climits>

#include "gstex_common.hpp"
#include "gstex_kernels.hpp"

namespace gstex {

constexpr int kSortThreads = 256;
// Items per thread per block tile: 16 (4096-item tiles) for long inputs, 8
// (2048-item tiles) below kSortSmallN, where 4096-item tiles would leave most
// of the 256 CUs idle.
constexpr int kSortSmallN = 1 << 22;
static inline int sort_items(int n) { return n < kSortSmallN ? 8 : 16; }
static inline int sort_blocks(int n) {
    const int tile = kSortThreads * sort_items(n);
    return (n + tile - 1) / tile;
}

// ------------------------------------------------------------------ fills
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
hipError_t fill(T* p, size_t n, T v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_fill<T><<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t fill_f32(float* p, size_t n, float v, hipStream_t s) { return fill<float>(p, n, v, s); }
hipError_t fill_i32(int* p, size_t n, int v, hipStream_t s) { return fill<int>(p, n, v, s); }
hipError_t fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    return fill<uint32_t>(p, n, v, s);
}

// ------------------------------------------------------------------ tile rectangle
// get_tile_bbox (helpers.cuh:37-51,75-92): [min, max) in tile units, clamped.
__device__ __forceinline__ void tile_rect(float cx, float cy, float ex, float ey, int bw,
                                          int tiles_x, int tiles_y, int& x0, int& y0, int& x1,
                                          int& y1) {
    const float fbw = (float)bw;
    const float tcx = cx / fbw, tcy = cy / fbw, tex = ex / fbw, tey = ey / fbw;
    x0 = min(max(0, f2i_trunc(tcx - tex)), tiles_x);
    x1 = min(max(0, f2i_trunc(tcx + tex + 1.f)), tiles_x);
    y0 = min(max(0, f2i_trunc(tcy - tey)), tiles_y);
    y1 = min(max(0, f2i_trunc(tcy + tey + 1.f)), tiles_y);
}

// get_tile_bbox_wrapped (helpers.cuh:53-73,94-111): the same rectangle
// unclamped, with the lower corner moved one tile further when it is <= 0;
// tile indices are then taken modulo the tile grid.
__device__ __forceinline__ void tile_rect_wrapped(float cx, float cy, float ex, float ey, int bw,
                                                  int& x0, int& y0, int& x1, int& y1) {
    const float fbw = (float)bw;
    const float tcx = cx / fbw, tcy = cy / fbw, tex = ex / fbw, tey = ey / fbw;
    x0 = f2i_trunc(tcx - tex);
    if (x0 <= 0 && x0 > INT_MIN) x0 -= 1;
    x1 = f2i_trunc(tcx + tex + 1.f);
    y0 = f2i_trunc(tcy - tey);
    if (y0 <= 0 && y0 > INT_MIN) y0 -= 1;
    y1 = f2i_trunc(tcy + tey + 1.f);
}

__device__ __forceinline__ int wrap_tile(int t, int n) {
    t %= n;
    return t < 0 ? t + n : t;
}

__device__ __forceinline__ bool writes_intersects(float ex, float ey) {
    // forward.cu:32 compares float extents with the double literal 1e-4.
    return !((double)ex <= 1e-4 && (double)ey <= 1e-4);
}

// ------------------------------------------------------------------ reference-layout keys
// One lane per Gaussian, writing (tile << 32 | depth bits, id) row-major over
// its tile rectangle from cum_tiles_hit[g-1]; outputs are pre-zeroed so slots
// a Gaussian counts but does not write stay (0, 0).  Writes past the
// Gaussian's own count window are dropped rather than spilling into the next
// Gaussian's slots (the walk stops there: a wrapped rectangle is unclamped).
__global__ __launch_bounds__(256) void k_map_intersects(
    int n, int num_isects, const float* __restrict__ centers, const float* __restrict__ extents,
    const float* __restrict__ depths, const int* __restrict__ cum, int tiles_x, int tiles_y,
    int bw, bool wrapped, int64_t* __restrict__ keys, int* __restrict__ ids) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float2 c = reinterpret_cast<const float2*>(centers)[g];
    const float2 e = reinterpret_cast<const float2*>(extents)[g];
    if (!writes_intersects(e.x, e.y)) return;
    int x0, y0, x1, y1;
    if (wrapped) tile_rect_wrapped(c.x, c.y, e.x, e.y, bw, x0, y0, x1, y1);
    else tile_rect(c.x, c.y, e.x, e.y, bw, tiles_x, tiles_y, x0, y0, x1, y1);
    int cur = g == 0 ? 0 : cum[g - 1];
    const int end = min(cum[g], num_isects);
    const int64_t did = (int64_t)__float_as_int(depths[g]);
    for (int ty = y0; ty < y1 && cur < end; ++ty)
        for (int tx = x0; tx < x1 && cur < end; ++tx) {
            const int ry = wrapped ? wrap_tile(ty, tiles_y) : ty;
            const int rx = wrapped ? wrap_tile(tx, tiles_x) : tx;
            if (cur >= 0) {
                keys[cur] = ((int64_t)(ry * tiles_x + rx) << 32) | did;
                ids[cur] = g;
            }
            ++cur;
        }
}

// Tile count of the wrapped rectangle (0 for a Gaussian that writes nothing):
// the count that makes wrapped binning write every tile of its rectangle.
__global__ __launch_bounds__(256) void k_num_tiles_hit_wrapped(int n,
                                                               const float* __restrict__ centers,
                                                               const float* __restrict__ extents,
                                                               int bw, int* __restrict__ out) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float2 c = reinterpret_cast<const float2*>(centers)[g];
    const float2 e = reinterpret_cast<const float2*>(extents)[g];
    int cnt = 0;
    if (writes_intersects(e.x, e.y)) {
        int x0, y0, x1, y1;
        tile_rect_wrapped(c.x, c.y, e.x, e.y, bw, x0, y0, x1, y1);
        const long long a = (long long)max(x1 - x0, 0) * (long long)max(y1 - y0, 0);
        cnt = (int)min(a, 0x7fffffffll);
    }
    out[g] = cnt;
}

hipError_t num_tiles_hit_wrapped(int n, const float* centers, const float* extents, int bw,
                                 int* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_num_tiles_hit_wrapped<<<(n + 255) / 256, 256, 0, s>>>(n, centers, extents, bw, out);
    return hipGetLastError();
}

// forward.cu:76-98 on sorted 64-bit keys (tile in the high word).
__global__ __launch_bounds__(256) void k_bin_edges64(int m, const int64_t* __restrict__ keys,
                                                     int num_tiles, int* __restrict__ bins) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int cur = (int)(keys[i] >> 32);
    const bool cur_ok = cur >= 0 && cur < num_tiles;
    if (i == 0 && cur_ok) bins[2 * cur] = 0;
    if (i == m - 1 && cur_ok) bins[2 * cur + 1] = m;
    if (i == 0) return;
    const int prev = (int)(keys[i - 1] >> 32);
    if (prev != cur) {
        if (prev >= 0 && prev < num_tiles) bins[2 * prev + 1] = i;
        if (cur_ok) bins[2 * cur] = i;
    }
}

// Same edges on sorted 32-bit tile ids, four per thread (one 16-byte load
// when the ids are 16-byte aligned, plus the id before them).  With id_bits
// > 0 the keys are packed (tile << id_bits | Gaussian id): the tile is the
// high part, and the ids go out to ids_out.
__global__ __launch_bounds__(256) void k_bin_edges32(int m, const uint32_t* __restrict__ tiles,
                                                     int num_tiles, int* __restrict__ bins,
                                                     bool vec, int id_bits,
                                                     int* __restrict__ ids_out) {
    const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= m) return;
    uint32_t raw[4];
    const bool full = vec && i0 + 3 < m;
    if (full) {
        const uint4 v = *reinterpret_cast<const uint4*>(tiles + i0);
        raw[0] = v.x; raw[1] = v.y; raw[2] = v.z; raw[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) raw[k] = i0 + k < m ? tiles[i0 + k] : 0u;
    }
    int cur[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) cur[k] = i0 + k < m ? (int)(raw[k] >> id_bits) : -1;
    if (id_bits > 0) {
        const uint32_t idm = (1u << id_bits) - 1u;
        if (full) {
            *reinterpret_cast<int4*>(ids_out + i0) =
                make_int4((int)(raw[0] & idm), (int)(raw[1] & idm), (int)(raw[2] & idm),
                          (int)(raw[3] & idm));
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (i0 + k < m) ids_out[i0 + k] = (int)(raw[k] & idm);
        }
    }
    int prev = i0 > 0 ? (int)(tiles[i0 - 1] >> id_bits) : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = i0 + k;
        if (i >= m) break;
        const int c = cur[k];
        const bool c_ok = c >= 0 && c < num_tiles;
        if (i == 0 && c_ok) bins[2 * c] = 0;
        if (i == m - 1 && c_ok) bins[2 * c + 1] = m;
        if (i > 0 && prev != c) {
            if (prev >= 0 && prev < num_tiles) bins[2 * prev + 1] = i;
            if (c_ok) bins[2 * c] = i;
        }
        prev = c;
    }
}

// ------------------------------------------------------------------ radix sort
template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// The lanes of the wave that hold the same digit d (of up to 8 bits) as this
// one, among the valid lanes: one ballot per digit bit, each lane keeping the
// lanes that agree with its own bit.  With t all ones when the lane's bit is
// set, the lanes that agree are ~(ballot ^ t): peers & ~(ballot ^ t) is one
// v_bitop3 per 32-bit half (truth table 0x90 over (peers, ballot, t)).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid, int bits) {
    const uint64_t v = __ballot(valid);
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (b >= bits) break;
        const uint32_t t = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
        const uint64_t bal = __ballot((int32_t)t < 0);
        lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)bal, t, 0x90);
        hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(bal >> 32), t, 0x90);
    }
    return ((uint64_t)hi << 32) | lo;
}

// Set bits of m in the lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Per-block digit histogram of one block tile (ITEMS * 256 items), stored
// digit-major (hist[d * nblocks + b]) so one exclusive scan yields scatter
// bases.  Block 0 also zeroes `nzero` words at `zero`.
template <typename K, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(int n, const K* __restrict__ keys,
                                                             K flip, int shift, int bits,
                                                             uint32_t* __restrict__ hist,
                                                             int nblocks,
                                                             uint64_t* __restrict__ zero,
                                                             int nzero) {
    // 16 copies of the counts (one per wave and lane quarter, 257 words
    // apart so that a digit's copies sit in different banks): lanes that
    // meet on one digit mostly add into different words
    constexpr int kCopies = 16, kStride = 257;
    __shared__ uint32_t cnt_all[kCopies * kStride];
    const int tid = threadIdx.x;
    uint32_t* const cnt = cnt_all + (tid >> 4) * kStride;
    // block 0 also clears the look-back states of the scan that follows
    if (blockIdx.x == 0)
        for (int i = tid; i < nzero; i += kSortThreads) zero[i] = 0ull;
    for (int i = tid; i < kCopies * kStride; i += kSortThreads) cnt_all[i] = 0;
    __syncthreads();
    const uint32_t mask = (1u << bits) - 1u;
    constexpr int kTile = kSortThreads * ITEMS;
    constexpr int kVec = 16 / sizeof(K);  // keys per 16-byte load
    const size_t base = (size_t)blockIdx.x * kTile;
    if (base + kTile <= (size_t)n && (reinterpret_cast<uintptr_t>(keys) & 15) == 0) {
        // a full tile: every load issued up front, 16 bytes per lane
        using V = uint4;
        V v[ITEMS / kVec];
        const V* kv = reinterpret_cast<const V*>(keys + base);
#pragma unroll
        for (int q = 0; q < ITEMS / kVec; ++q) v[q] = kv[q * kSortThreads + tid];
#pragma unroll
        for (int q = 0; q < ITEMS / kVec; ++q) {
            const K* k = reinterpret_cast<const K*>(&v[q]);
#pragma unroll
            for (int j = 0; j < kVec; ++j)
                atomicAdd(&cnt[digit_of(k[j] ^ flip, shift, mask)], 1u);
        }
    } else {
#pragma unroll 4
        for (int r = 0; r < ITEMS; ++r) {
            const size_t idx = base + (size_t)r * kSortThreads + tid;
            if (idx < (size_t)n)
                atomicAdd(&cnt[digit_of(keys[idx] ^ flip, shift, mask)], 1u);
        }
    }
    __syncthreads();
    if (tid < (1 << bits)) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kCopies; ++w) c += cnt_all[w * kStride + tid];
        hist[(size_t)tid * nblocks + blockIdx.x] = c;
    }
}

// Exclusive scan of up to 64K uint32 values in place by one 1024-lane
// workgroup.  Wave w owns the PER * 64 contiguous values from w * PER * 64 on,
// as PER / 4 chunks of 256: in chunk q lane l holds values q * 256 + 4 l ..
// + 3, so each 16-byte load or store instruction covers 1 KB of contiguous
// memory.  All loads are issued before any value is used (the scan costs about
// two memory latencies).  Wave totals are exchanged through LDS; each chunk
// is then scanned within the wave (four values per lane, a shuffle scan of the
// lane sums) on top of a running carry, and written back from registers.  vec
// is false when a is not 16-byte aligned (scalar accesses only).
template <int PER>
__global__ __launch_bounds__(1024) void k_scan_single(uint32_t* __restrict__ a, int n,
                                                      uint32_t* __restrict__ total, bool vec) {
    __shared__ uint32_t wave_sums[1024 / kWave];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int seg = wave * PER * kWave + 4 * lane;
    uint32_t v[PER];
#pragma unroll
    for (int q = 0; q < PER / 4; ++q) {
        const int i = seg + q * 256;
        if (vec && i + 3 < n) {
            const uint4 x = *reinterpret_cast<const uint4*>(a + i);
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[4 * q + j] = i + j < n ? a[i + j] : 0u;
        }
    }
    uint32_t lsum = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) lsum += v[j];
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) wave_sums[wave] = lsum;
    __syncthreads();
    uint32_t carry = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 1024 / kWave; ++w) {
        const uint32_t t = wave_sums[w];
        if (w < wave) carry += t;
        tot += t;
    }
#pragma unroll
    for (int q = 0; q < PER / 4; ++q) {
        const uint32_t q0 = v[4 * q], q1 = q0 + v[4 * q + 1], q2 = q1 + v[4 * q + 2],
                       q3 = q2 + v[4 * q + 3];
        uint32_t inc = q3;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t t = __shfl_up(inc, off);
            if (lane >= off) inc += t;
        }
        const uint32_t b0 = carry + inc - q3;
        carry += __shfl(inc, kWave - 1);
        const uint32_t o[4] = {b0, b0 + q0, b0 + q1, b0 + q2};
        const int i = seg + q * 256;
        if (vec && i + 3 < n) {
            *reinterpret_cast<uint4*>(a + i) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (i + j < n) a[i + j] = o[j];
        }
    }
    if (tid == 0 && total) *total = tot;
}

// One-workgroup scan of n <= kScanSingleMax values, with the per-lane segment
// length the smallest that covers n.
static void launch_scan_single(uint32_t* a, int n, uint32_t* total, hipStream_t s) {
    const bool vec = ((uintptr_t)a & 15u) == 0;
    if (n <= 4 * 1024) k_scan_single<4><<<1, 1024, 0, s>>>(a, n, total, vec);
    else if (n <= 8 * 1024) k_scan_single<8><<<1, 1024, 0, s>>>(a, n, total, vec);
    else if (n <= 16 * 1024) k_scan_single<16><<<1, 1024, 0, s>>>(a, n, total, vec);
    else if (n <= 32 * 1024) k_scan_single<32><<<1, 1024, 0, s>>>(a, n, total, vec);
    else k_scan_single<64><<<1, 1024, 0, s>>>(a, n, total, vec);
}

// Device-wide exclusive scan in three passes: per-tile sums, a scan of the
// tile sums, then each tile scans itself from its offset.  A tile is 4096
// values, 16 contiguous per lane.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanThreads * kScanPer;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves,
                                                         uint32_t& block_total) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t t = __shfl_up(inc, off);
        if (lane >= off) inc += t;
    }
    if (lane == kWave - 1) lds_waves[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; ++w) {
        const uint32_t s = lds_waves[w];
        if (w < wave) wave_off += s;
        tot += s;
    }
    block_total = tot;
    return wave_off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const uint32_t* __restrict__ a,
                                                                 int n,
                                                                 uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[kScanThreads / kWave];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j)
        if (base + j < (size_t)n) s += a[base + j];
    uint32_t tot;
    block_exclusive_scan(s, lds, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile_apply(uint32_t* __restrict__ a, int n,
                                                                  const uint32_t* __restrict__ offs) {
    __shared__ uint32_t lds[kScanThreads / kWave];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        v[j] = base + j < (size_t)n ? a[base + j] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = offs[blockIdx.x] + block_exclusive_scan(s, lds, tot);
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        if (base + j < (size_t)n) a[base + j] = run;
        run += v[j];
    }
}

// Up to this many values one 1024-lane workgroup scans faster than three launches.
constexpr int kScanSingleMax = 1 << 16;
static size_t scan_partials_count(int n) { return (size_t)(n + kScanTile - 1) / kScanTile + 1; }

// In-place exclusive scan of n uint32 values; writes the total to *total if
// given.  partials holds scan_partials_count(n) values.
static hipError_t scan_exclusive(uint32_t* a, int n, uint32_t* total, uint32_t* partials,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n <= kScanSingleMax) {
        launch_scan_single(a, n, total, s);
        return hipGetLastError();
    }
    const int tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles > kScanSingleMax) return hipErrorInvalidValue;
    k_scan_tile_sums<<<tiles, kScanThreads, 0, s>>>(a, n, partials);
    launch_scan_single(partials, tiles, total, s);
    k_scan_tile_apply<<<tiles, kScanThreads, 0, s>>>(a, n, partials);
    return hipGetLastError();
}

// Stable scatter of one block tile (ITEMS * 256 items).  Wave w reads the
// ITEMS * 64 consecutive items from w * ITEMS * 64 on (item r of lane l is
// number w * ITEMS * 64 + r * 64 + l, one coalesced read per round), so the
// order (wave, round, lane) is the input order.  Ranking needs no barrier per
// round: each wave ranks its items round by round with ballots over the digit
// bits against a wave-private running count per digit in LDS (all lanes read
// the count, then the leader of each digit group adds the group size; one
// wave's LDS operations complete in program order).  One block-wide pass then
// turns the per-wave counts into each wave's start inside each digit's run of
// the tile, the items are placed in an LDS copy of the tile sorted by digit,
// and lanes write consecutive LDS positions out, so each digit's run goes to
// consecutive global addresses (hist[digit][block] onwards).
struct ScatterLds {
    uint32_t gbase[256];                      // global start of each digit's run for this block
    uint32_t lstart[256];                     // start of each digit's run in the LDS tile
    uint32_t wrun[kSortThreads / kWave][256]; // per wave: running count, then start offset
};

// The rank and place part of the scatter, for a tile already held in
// registers in (wave, round, lane) order: item r of lane l of wave w is item
// w * ITEMS * 64 + r * 64 + l of the tile, and items at or past count are
// absent.  skeys / svals are LDS arrays of the tile's size (svals unused
// without values); gb is thread t's digit base (the global start of digit
// t's run for this tile, 0 for t past the digits), loaded by the caller
// ahead of time.  On return skeys holds the tile sorted stably by digit,
// sl.lstart[d] is where digit d's run starts in it and sl.gbase[d] its
// global start (all threads synchronised).
template <typename K, int ITEMS>
__device__ __forceinline__ void tile_rank_place(const K (&key)[ITEMS],
                                                const uint32_t (&val)[ITEMS], int count,
                                                bool has_vals, K flip, int shift, int bits,
                                                uint32_t gb, ScatterLds& sl, K* skeys,
                                                uint32_t* svals) {
    constexpr int kWaves = kSortThreads / kWave;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const uint32_t mask = (1u << bits) - 1u;
    const int wbase = wave * ITEMS * kWave;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) sl.wrun[w][tid] = 0;
    __syncthreads();

    // wave-private ranking: rank[r] = items of digit d before item r in this wave
    uint32_t rank[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const bool valid = wbase + r * kWave + lane < count;
        const uint32_t d = valid ? digit_of(key[r] ^ flip, shift, mask) : 0u;
        const uint64_t peers = digit_peers(d, valid, bits);
        const uint32_t before = lanes_below(peers);
        const uint32_t seen = valid ? sl.wrun[wave][d] : 0u;
        if (valid && before == 0) sl.wrun[wave][d] = seen + (uint32_t)__popcll(peers);
        rank[r] = seen + before;
    }
    __syncthreads();
    // per digit (one per thread): each wave's start inside the digit's run,
    // and the exclusive scan of the digit totals over the tile
    {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t t = sl.wrun[w][tid];
            sl.wrun[w][tid] = c;
            c += t;
        }
        uint32_t v = c;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t t = __shfl_up(v, off);
            if (lane >= off) v += t;
        }
        sl.lstart[tid] = v;  // inclusive within the wave for now
        sl.gbase[tid] = gb;
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < wave; ++w) add += sl.lstart[w * kWave + kWave - 1];
        __syncthreads();
        sl.lstart[tid] = v + add - c;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        if (wbase + r * kWave + lane < count) {
            const uint32_t d = digit_of(key[r] ^ flip, shift, mask);
            const uint32_t pos = sl.lstart[d] + sl.wrun[wave][d] + rank[r];
            skeys[pos] = key[r];
            if (has_vals) svals[pos] = val[r];
        }
    }
    __syncthreads();
}

// The rank / place / write part of the scatter (tile_rank_place, then each
// digit's run written to consecutive global addresses); writes stay below
// out_n.
template <typename K, int ITEMS>
__device__ __forceinline__ void tile_rank_scatter(const K (&key)[ITEMS],
                                                  const uint32_t (&val)[ITEMS], int count,
                                                  bool has_vals, K flip, int shift, int bits,
                                                  uint32_t gb, K* __restrict__ keys_out,
                                                  uint32_t* __restrict__ vals_out, uint32_t out_n,
                                                  ScatterLds& sl, K* skeys,
                                                  uint32_t* svals) {
    tile_rank_place<K, ITEMS>(key, val, count, has_vals, flip, shift, bits, gb, sl, skeys,
                              svals);
    const int tid = threadIdx.x;
    const uint32_t mask = (1u << bits) - 1u;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int p = r * kSortThreads + tid;
        if (p < count) {
            const K k = skeys[p];
            const uint32_t d = digit_of(k ^ flip, shift, mask);
            const uint32_t gp = sl.gbase[d] + (uint32_t)p - sl.lstart[d];
            if (gp < out_n) {  // always, when the digit counts are those of the keys
                keys_out[gp] = k;
                if (has_vals) vals_out[gp] = svals[p];
            }
        }
    }
}

// VALS false: keys only (vals_in / vals_out unused), and no LDS for values.
template <typename K, int ITEMS, bool VALS>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    int n, const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    K* __restrict__ keys_out, uint32_t* __restrict__ vals_out, K flip, int shift, int bits,
    const uint32_t* __restrict__ hist, int nblocks) {
    constexpr int kTile = kSortThreads * ITEMS;
    __shared__ ScatterLds sl;
    __shared__ K skeys[kTile];
    __shared__ uint32_t svals[VALS ? kTile : 1];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const size_t base = (size_t)blockIdx.x * kTile;
    const int count = (int)min((size_t)kTile, (size_t)n - base);
    const int wbase = wave * ITEMS * kWave;

    K key[ITEMS];
    uint32_t val[ITEMS];
    constexpr bool has_vals = VALS;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int li = wbase + r * kWave + lane;
        key[r] = li < count ? keys_in[base + li] : (K)0;
        val[r] = has_vals && li < count ? vals_in[base + li] : 0u;
    }
    const uint32_t gb = tid < (1 << bits) ? hist[(size_t)tid * nblocks + blockIdx.x] : 0u;
    tile_rank_scatter<K, ITEMS>(key, val, count, has_vals, flip, shift, bits, gb, keys_out,
                                vals_out, (uint32_t)n, sl, skeys, svals);
}

// The last tile-sort pass of packed binning keys (tile << id_bits | id), by
// digit [shift, shift + bits): the stable scatter of k_radix_scatter (keys
// only), writing the ids (key & id mask) at the sorted positions and the bin
// edges (bins[2 t] first position of tile t, bins[2 t + 1] one past its
// last; tiles >= num_tiles belong to no bin) of every tile change inside one
// digit run of the tile, from the LDS-sorted tile.  The keys are written only
// at the two ends of each run, for k_run_edges, which settles the edges where
// runs meet.
template <int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_last_pass_edges(
    int n, const uint32_t* __restrict__ keys_in, uint32_t* __restrict__ keys_out,
    int* __restrict__ ids_out, int id_bits, int shift, int bits,
    const uint32_t* __restrict__ hist, int nblocks, int num_tiles, int* __restrict__ bins) {
    constexpr int kTile = kSortThreads * ITEMS;
    __shared__ ScatterLds sl;
    __shared__ uint32_t skeys[kTile];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const size_t base = (size_t)blockIdx.x * kTile;
    const int count = (int)min((size_t)kTile, (size_t)n - base);
    const int wbase = wave * ITEMS * kWave;
    uint32_t key[ITEMS], val[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int li = wbase + r * kWave + lane;
        key[r] = li < count ? keys_in[base + li] : 0u;
        val[r] = 0u;
    }
    const uint32_t gb = tid < (1 << bits) ? hist[(size_t)tid * nblocks + blockIdx.x] : 0u;
    tile_rank_place<uint32_t, ITEMS>(key, val, count, false, 0u, shift, bits, gb, sl, skeys,
                                     nullptr);
    const uint32_t mask = (1u << bits) - 1u, idmask = (1u << id_bits) - 1u;
    const uint32_t nt = (uint32_t)num_tiles;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int p = r * kSortThreads + tid;
        if (p < count) {
            const uint32_t k = skeys[p];
            const uint32_t d = (k >> shift) & mask;
            const uint32_t gp = sl.gbase[d] + (uint32_t)p - sl.lstart[d];
            if (gp < (uint32_t)n) {  // always, when the digit counts are those of the keys
                ids_out[gp] = (int)(k & idmask);
                const uint32_t t = k >> id_bits;
                // the run of digit d ends where the next digit's starts
                const bool first = (uint32_t)p == sl.lstart[d];
                const bool last = (uint32_t)(p + 1) == (d < mask ? sl.lstart[d + 1] : (uint32_t)count);
                if (first || last) keys_out[gp] = k;
                if (!first) {
                    const uint32_t tp = skeys[p - 1] >> id_bits;
                    if (tp != t) {
                        if (tp < nt) bins[2 * tp + 1] = (int)gp;
                        if (t < nt) bins[2 * t] = (int)gp;
                    }
                } else if (gp == 0 && t < nt) {
                    bins[2 * t] = 0;
                }
                if (gp + 1 == (uint32_t)n && t < nt) bins[2 * t + 1] = n;
            }
        }
    }
}

// The bin edges where two runs of k_last_pass_edges meet: run r (digit-major,
// r = d * nblocks + b) starts at hist[r] (the scanned digit counts) and ends
// where run r + 1 starts; its first key and the previous run's last key were
// written.
__global__ __launch_bounds__(256) void k_run_edges(int n, int runs,
                                                   const uint32_t* __restrict__ hist,
                                                   const uint32_t* __restrict__ keys, int id_bits,
                                                   int num_tiles, int* __restrict__ bins) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    const uint32_t pos = hist[r], end = r + 1 < runs ? hist[r + 1] : (uint32_t)n;
    if (end <= pos || pos == 0 || pos >= (uint32_t)n) return;
    const uint32_t t = keys[pos] >> id_bits, tp = keys[pos - 1] >> id_bits;
    const uint32_t nt = (uint32_t)num_tiles;
    if (t != tp) {
        if (tp < nt) bins[2 * tp + 1] = (int)pos;
        if (t < nt) bins[2 * t] = (int)pos;
    }
}

// ------------------------------------------------------------------ one-sweep sort
// LSD radix sort in which every pass is one launch: the digit counts of all
// passes come from one read of the keys beforehand (k_onesweep_hist), and each
// block finds where its digits start by a decoupled look-back over the
// blocks before it (k_onesweep_scatter).  A block's tile is the one its
// ticket names (an atomic counter taken when the block starts), so a block
// only ever waits for blocks that are already running; the wait ends when
// the block before it has published its inclusive count.
//
// Look-back state of one (block, digit), 64 bits: status in bits 62-63
// (0 not yet written, aggregate = this block's count, inclusive = the count
// of this block and every earlier one), count in bits 0-61.
constexpr uint64_t kLbAggregate = 1ull << 62;
constexpr uint64_t kLbInclusive = 2ull << 62;
constexpr uint64_t kLbCount = (1ull << 62) - 1ull;
// Bound on the polls one thread makes over a whole look-back.  A wait that
// ran this long would mean a block stopped without publishing: the thread
// then records it in *err and goes on with the partial sum, which can only
// be smaller than the true one, so every store stays in bounds.
constexpr int kLbPollLimit = 1 << 20;

__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the counts of blocks 0 .. blk-1 in the chain states[j * stride + d],
// for one thread: kLbWindow states are read per round trip (blocks j, j-1, ...)
// and consumed in order up to the first inclusive one, or up to the first
// not yet written, which is then polled again.  Block 0's states are
// inclusive, so the walk never passes it.
constexpr int kLbWindow = 8;
__device__ __forceinline__ uint64_t lb_exclusive(const uint64_t* states, int blk, int stride,
                                                 int d, uint32_t* err) {
    uint64_t sum = 0;
    int j = blk - 1, polls = 0;
    while (j >= 0) {
        uint64_t s[kLbWindow];
#pragma unroll
        for (int w = 0; w < kLbWindow; ++w)
            s[w] = j - w >= 0 ? lb_load(states + (size_t)(j - w) * stride + d) : kLbInclusive;
        int used = 0;
        bool done = false, stalled = false;
#pragma unroll
        for (int w = 0; w < kLbWindow; ++w) {
            if (done || stalled) continue;
            const uint64_t st = s[w] & ~kLbCount;
            if (st == 0) {
                stalled = true;
                continue;
            }
            sum += s[w] & kLbCount;
            ++used;
            done = st == kLbInclusive;
        }
        if (done) break;
        j -= used;
        if (stalled) {
            if (++polls > kLbPollLimit) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return sum;
}

// The same sum for a whole wave at once (one chain, stride 1): lane l reads
// block j - l, so one round trip covers 64 blocks.  Every lane returns the sum.
__device__ __forceinline__ uint64_t lb_exclusive_wave(const uint64_t* states, int blk,
                                                      uint32_t* err) {
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t sum = 0;
    int j = blk - 1, polls = 0;
    while (j >= 0) {
        const uint64_t s = j - lane >= 0 ? lb_load(states + (j - lane)) : kLbInclusive;
        const uint64_t st = s & ~kLbCount;
        const uint64_t not_ready = __ballot(st == 0), incl = __ballot(st == kLbInclusive);
        // lanes before the first unwritten state, through the first inclusive one
        const int first_nr = not_ready ? __ffsll((unsigned long long)not_ready) - 1 : kWave;
        const int first_in = incl ? __ffsll((unsigned long long)incl) - 1 : kWave;
        const int take = first_in < first_nr ? first_in + 1 : first_nr;
        uint64_t part = lane < take ? (s & kLbCount) : 0ull;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) part += __shfl_xor(part, off);
        sum += part;
        if (first_in < first_nr) break;
        j -= take;
        if (take < kWave) {
            if (++polls > kLbPollLimit) {
                if (lane == 0) atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return sum;
}

constexpr int kOsMaxPasses = 4;  // 32-bit keys, digits of up to 8 bits
constexpr int kOsMaxBlocks = 1024;  // block tiles up to which the one-sweep passes are used
constexpr int kOsHistItems = 2048;

// Global digit counts of every pass in one read: hist[p * 256 + d] += items
// whose pass-p digit is d (hist zeroed beforehand).
__global__ __launch_bounds__(256) void k_onesweep_hist(int n, const uint32_t* __restrict__ keys,
                                                       int passes, int start_bit, int bits_per,
                                                       int end_bit, int per_block,
                                                       uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kOsMaxPasses * 256];
    const int tid = threadIdx.x;
    for (int i = tid; i < passes * 256; i += 256) cnt[i] = 0;
    __syncthreads();
    const size_t lo = (size_t)blockIdx.x * per_block;
    const size_t hi = min(lo + (size_t)per_block, (size_t)n);
    for (size_t i = lo + tid; i < hi; i += 256) {
        const uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) {
            const int shift = start_bit + p * bits_per;
            const int bits = min(bits_per, end_bit - shift);
            atomicAdd(&cnt[p * 256 + ((k >> shift) & ((1u << bits) - 1u))], 1u);
        }
    }
    __syncthreads();
    for (int i = tid; i < passes * 256; i += 256) {
        const uint32_t c = cnt[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// One pass of the one-sweep sort over block tiles of ITEMS * 256 items.  The
// tile is read, ranked and placed in LDS as in k_radix_scatter; its digit
// counts are published before the look-back, and each digit's run goes to
// base(digit) + (items of the digit in earlier tiles) onwards.
template <int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_onesweep_scatter(
    int n, const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int shift, int bits,
    const uint32_t* __restrict__ digit_counts, uint64_t* __restrict__ states,
    uint32_t* __restrict__ ticket, uint32_t* __restrict__ err) {
    constexpr int kTile = kSortThreads * ITEMS;
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t lstart[256];
    __shared__ uint32_t wrun[kWaves][256];
    __shared__ uint32_t skeys[kTile];
    __shared__ uint32_t svals[kTile];
    __shared__ uint32_t s_blk;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const uint32_t mask = (1u << bits) - 1u;
    const int ndig = 1 << bits;
    if (tid == 0) s_blk = atomicAdd(ticket, 1u);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wrun[w][tid] = 0;
    __syncthreads();
    const int blk = (int)s_blk;
    const size_t base = (size_t)blk * kTile;
    const int count = (int)min((size_t)kTile, (size_t)n - base);
    const int wbase = wave * ITEMS * kWave;

    uint32_t key[ITEMS], val[ITEMS];
    const bool has_vals = vals_in != nullptr;  // keys only when null
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int li = wbase + r * kWave + lane;
        key[r] = li < count ? keys_in[base + li] : 0u;
        val[r] = has_vals && li < count ? vals_in[base + li] : 0u;
    }
    uint32_t rank[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const bool valid = wbase + r * kWave + lane < count;
        const uint32_t d = valid ? (key[r] >> shift) & mask : 0u;
        const uint64_t peers = digit_peers(d, valid, bits);
        const uint32_t before = lanes_below(peers);
        const uint32_t seen = valid ? wrun[wave][d] : 0u;
        if (valid && before == 0) wrun[wave][d] = seen + (uint32_t)__popcll(peers);
        rank[r] = seen + before;
    }
    __syncthreads();
    uint32_t c = 0, dbase;
    {
        // this tile's count of digit tid, published at once for later tiles
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t t = wrun[w][tid];
            wrun[w][tid] = c;
            c += t;
        }
        if (tid < ndig)
            lb_store(states + (size_t)blk * 256 + tid,
                     (blk == 0 ? kLbInclusive : kLbAggregate) | (uint64_t)c);
        // exclusive scans over the digits of the tile counts (LDS run
        // starts) and of the global counts (digit bases)
        const uint32_t g = tid < ndig ? digit_counts[tid] : 0u;
        uint32_t v = c, gv = g;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t t = __shfl_up(v, off), tg = __shfl_up(gv, off);
            if (lane >= off) {
                v += t;
                gv += tg;
            }
        }
        lstart[tid] = v;
        gbase[tid] = gv;
        __syncthreads();
        uint32_t add = 0, gadd = 0;
        for (int w = 0; w < wave; ++w) {
            add += lstart[w * kWave + kWave - 1];
            gadd += gbase[w * kWave + kWave - 1];
        }
        __syncthreads();
        lstart[tid] = v + add - c;
        dbase = gv + gadd - g;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        if (wbase + r * kWave + lane < count) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t pos = lstart[d] + wrun[wave][d] + rank[r];
            skeys[pos] = key[r];
            if (has_vals) svals[pos] = val[r];
        }
    }
    if (tid < ndig) {
        uint64_t before = 0;
        if (blk > 0) {
            before = lb_exclusive(states, blk, 256, tid, err);
            lb_store(states + (size_t)blk * 256 + tid, kLbInclusive | (before + c));
        }
        gbase[tid] = dbase + (uint32_t)before;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int p = r * kSortThreads + tid;
        if (p < count) {
            const uint32_t k = skeys[p];
            const uint32_t d = (k >> shift) & mask;
            const uint32_t gp = gbase[d] + (uint32_t)p - lstart[d];
            keys_out[gp] = k;
            if (has_vals) vals_out[gp] = svals[p];
        }
    }
}

// Tiles of the single-pass look-back scans: kWrScanPer consecutive values per
// thread, 256 threads.
constexpr int kWrScanPer = 16;
constexpr int kWrScanTile = 256 * kWrScanPer;
static int wr_scan_tiles(int n) { return (n + kWrScanTile - 1) / kWrScanTile; }

// In-place exclusive scan of n uint32 values in one launch: tiles of
// kWrScanTile values, 16 consecutive per thread, each tile adding the sum of
// the tiles before it by a wave-wide decoupled look-back (ticket order).
// states: one look-back state per tile, then the ticket and error words, all
// zeroed beforehand.
__global__ __launch_bounds__(256) void k_scan_lookback(uint32_t* __restrict__ a, int n,
                                                       uint64_t* __restrict__ states, int tiles) {
    __shared__ uint32_t lds[256 / kWave];
    __shared__ uint32_t s_blk, s_before;
    uint32_t* const ticket = (uint32_t*)(states + tiles);
    const int tid = threadIdx.x;
    if (tid == 0) s_blk = atomicAdd(ticket, 1u);
    __syncthreads();
    const int blk = (int)s_blk;
    const size_t base = (size_t)blk * kWrScanTile + (size_t)tid * kWrScanPer;
    uint32_t v[kWrScanPer];
    uint32_t sum = 0;
    if (base + kWrScanPer <= (size_t)n) {
#pragma unroll
        for (int q = 0; q < kWrScanPer / 4; ++q) {
            const uint4 x = reinterpret_cast<const uint4*>(a + base)[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kWrScanPer; ++j) v[j] = base + j < (size_t)n ? a[base + j] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kWrScanPer; ++j) sum += v[j];
    uint32_t tot;
    const uint32_t excl = block_exclusive_scan(sum, lds, tot);
    if (tid < kWave) {
        if (tid == 0)
            lb_store(states + blk, (blk == 0 ? kLbInclusive : kLbAggregate) | (uint64_t)tot);
        uint64_t before = 0;
        if (blk > 0) {
            before = lb_exclusive_wave(states, blk, ticket + 1);
            if (tid == 0) lb_store(states + blk, kLbInclusive | (before + tot));
        }
        if (tid == 0) s_before = (uint32_t)before;
    }
    __syncthreads();
    uint32_t run = s_before + excl;
    if (base + kWrScanPer <= (size_t)n) {
#pragma unroll
        for (int q = 0; q < kWrScanPer / 4; ++q) {
            const uint32_t o0 = run, o1 = o0 + v[4 * q], o2 = o1 + v[4 * q + 1],
                           o3 = o2 + v[4 * q + 2];
            reinterpret_cast<uint4*>(a + base)[q] = make_uint4(o0, o1, o2, o3);
            run = o3 + v[4 * q + 3];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kWrScanPer; ++j) {
            if (base + j < (size_t)n) a[base + j] = run;
            run += v[j];
        }
    }
}

// Workspace of the one-sweep sort after the key and value buffers, in 32-bit
// words: digit counts (kOsMaxPasses x 256), kOsMaxPasses tickets and the
// error word (padded to kOsHead words), then the look-back states
// (kOsMaxPasses x blocks x 256, 8 bytes each), so that one fill clears
// everything a sort uses.
constexpr int kOsHead = kOsMaxPasses * 256 + 16;
static size_t onesweep_state_words(int n) {
    if (sort_blocks(n) > kOsMaxBlocks) return 0;
    return (size_t)kOsHead + 2 * (size_t)kOsMaxPasses * sort_blocks(n) * 256;
}

template <typename K>
size_t radix_workspace_bytes(int n) {
    const size_t nb = (size_t)sort_blocks(n);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t lsd = up(sizeof(uint32_t) * 256 * nb) +
                       up(sizeof(uint64_t) * scan_partials_count((int)(256 * nb)));
    const size_t os = sizeof(K) == 4 ? up(sizeof(uint32_t) * onesweep_state_words(n)) : 0;
    return up(sizeof(K) * (size_t)n) + up(sizeof(uint32_t) * (size_t)n) + std::max(lsd, os) + 256;
}

// One-sweep sort of (uint32 key, uint32 value) pairs over key bits [0, end_bit).
static hipError_t onesweep_sort_u32(int n, const uint32_t* keys_in, const uint32_t* vals_in,
                                    uint32_t* keys_out, uint32_t* vals_out, int start_bit,
                                    int end_bit, uint32_t* tk, uint32_t* tv, uint32_t* state,
                                    hipStream_t s) {
    const int passes = (end_bit - start_bit + 7) / 8;
    const int bits_per = (end_bit - start_bit + passes - 1) / passes;
    const int nblocks = sort_blocks(n);
    const bool small = sort_items(n) == 8;
    uint32_t* hist = state;
    uint32_t* tickets = state + kOsMaxPasses * 256;
    uint32_t* err = tickets + kOsMaxPasses;
    uint64_t* states = (uint64_t*)(state + kOsHead);
    hipError_t e = fill<uint32_t>(state, (size_t)kOsHead + 2 * (size_t)passes * nblocks * 256,
                                  0u, s);
    if (e != hipSuccess) return e;
    const int hb = std::min((n + kOsHistItems - 1) / kOsHistItems, 1024);
    const int per = (n + hb - 1) / hb;
    k_onesweep_hist<<<hb, 256, 0, s>>>(n, keys_in, passes, start_bit, bits_per, end_bit, per,
                                       hist);
    const uint32_t* sk = keys_in;
    const uint32_t* sv = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = start_bit + pass * bits_per;
        const int bits = std::min(bits_per, end_bit - shift);
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        uint32_t* dk = to_out ? keys_out : tk;
        uint32_t* dv = vals_in ? (to_out ? vals_out : tv) : nullptr;
        uint64_t* st = states + (size_t)pass * nblocks * 256;
        if (small)
            k_onesweep_scatter<8><<<nblocks, kSortThreads, 0, s>>>(
                n, sk, sv, dk, dv, shift, bits, hist + pass * 256, st, tickets + pass, err);
        else
            k_onesweep_scatter<16><<<nblocks, kSortThreads, 0, s>>>(
                n, sk, sv, dk, dv, shift, bits, hist + pass * 256, st, tickets + pass, err);
        sk = dk;
        sv = dv;
    }
    return hipGetLastError();
}

// Stable LSD sort of (key, value) pairs over key bits [start_bit, end_bit)
// (bits outside are carried); with vals_in null the keys are sorted alone.
// With all 64 bits of a 64-bit key the sign bit is flipped in the digits,
// giving the signed (int64) order.
template <typename K>
hipError_t radix_sort_pairs(int n, const K* keys_in, const uint32_t* vals_in, K* keys_out,
                            uint32_t* vals_out, int end_bit, void* ws, size_t ws_bytes,
                            hipStream_t s, int start_bit = 0) {
    if (n <= 0) return hipSuccess;
    if (ws_bytes < radix_workspace_bytes<K>(n)) return hipErrorInvalidValue;
    const int nblocks = sort_blocks(n);
    const bool small = sort_items(n) == 8;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char* p = (char*)ws;
    K* tk = (K*)p;
    p += up(sizeof(K) * (size_t)n);
    uint32_t* tv = (uint32_t*)p;
    p += up(sizeof(uint32_t) * (size_t)n);
    uint32_t* hist = (uint32_t*)p;
    p += up(sizeof(uint32_t) * 256 * (size_t)nblocks);
    uint32_t* partials = (uint32_t*)p;
    const int passes = (end_bit - start_bit + 7) / 8;
    if (passes <= 0) {
        hipError_t e = hipMemcpyAsync(keys_out, keys_in, sizeof(K) * (size_t)n,
                                      hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess || !vals_in) return e;
        return hipMemcpyAsync(vals_out, vals_in, sizeof(uint32_t) * (size_t)n,
                              hipMemcpyDeviceToDevice, s);
    }
    // The one-sweep passes win while the look-back chains are short; over many
    // block tiles the serialised round trips between a block and the ones
    // before it (coherent across the chip's L2s) cost more than a separate
    // histogram and scan.  Measured at 1080p / 500k: the 500k-key depth sort
    // (245 tiles) 108 -> 78 us, the 11.5M-key tile sort (2805 tiles) 173 -> 197 us.
    if constexpr (sizeof(K) == 4)
        if (nblocks <= kOsMaxBlocks)
            return onesweep_sort_u32(n, (const uint32_t*)keys_in, vals_in, (uint32_t*)keys_out,
                                     vals_out, start_bit, end_bit, (uint32_t*)tk, tv, hist, s);
    const int bits_per = (end_bit - start_bit + passes - 1) / passes;  // balanced digit widths <= 8
    const K flip = (sizeof(K) == 8 && end_bit == 64) ? (K)((K)1 << (8 * sizeof(K) - 1)) : (K)0;
    const K* sk = keys_in;
    const uint32_t* sv = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = start_bit + pass * bits_per;
        const int bits = min(bits_per, end_bit - shift);
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        K* dk = to_out ? keys_out : tk;
        uint32_t* dv = vals_in ? (to_out ? vals_out : tv) : nullptr;
        // past one workgroup's reach, the digit counts are scanned in one
        // launch with a look-back (its states cleared by the histogram kernel)
        const int hcount = (1 << bits) * nblocks;
        const bool lb = hcount > kScanSingleMax;
        const int lbt = wr_scan_tiles(hcount);
        uint64_t* lbs = lb ? reinterpret_cast<uint64_t*>(partials) : nullptr;
        if (small)
            k_radix_hist<K, 8><<<nblocks, kSortThreads, 0, s>>>(n, sk, flip, shift, bits, hist,
                                                                 nblocks, lbs, lb ? lbt + 1 : 0);
        else
            k_radix_hist<K, 16><<<nblocks, kSortThreads, 0, s>>>(n, sk, flip, shift, bits, hist,
                                                                  nblocks, lbs, lb ? lbt + 1 : 0);
        if (lb) {
            k_scan_lookback<<<lbt, 256, 0, s>>>(hist, hcount, lbs, lbt);
        } else {
            hipError_t es = scan_exclusive(hist, hcount, nullptr, partials, s);
            if (es != hipSuccess) return es;
        }
        if (small && sv)
            k_radix_scatter<K, 8, true><<<nblocks, kSortThreads, 0, s>>>(
                n, sk, sv, dk, dv, flip, shift, bits, hist, nblocks);
        else if (small)
            k_radix_scatter<K, 8, false><<<nblocks, kSortThreads, 0, s>>>(
                n, sk, sv, dk, dv, flip, shift, bits, hist, nblocks);
        else if (sv)
            k_radix_scatter<K, 16, true><<<nblocks, kSortThreads, 0, s>>>(
                n, sk, sv, dk, dv, flip, shift, bits, hist, nblocks);
        else
            k_radix_scatter<K, 16, false><<<nblocks, kSortThreads, 0, s>>>(
                n, sk, sv, dk, dv, flip, shift, bits, hist, nblocks);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        sk = dk;
        sv = dv;
    }
    return hipSuccess;
}

size_t sort_u32_pairs_workspace_bytes(int n) { return radix_workspace_bytes<uint32_t>(n); }

hipError_t sort_u32_pairs(int n, const uint32_t* keys_in, const uint32_t* vals_in,
                          uint32_t* keys_out, uint32_t* vals_out, int end_bit, void* ws,
                          size_t ws_bytes, hipStream_t s) {
    return radix_sort_pairs<uint32_t>(n, keys_in, vals_in, keys_out, vals_out, end_bit, ws,
                                      ws_bytes, s);
}

size_t sort_isects_workspace_bytes(int n) { return radix_workspace_bytes<uint64_t>(n); }

hipError_t sort_isects(int n, const int64_t* keys_in, const int* ids_in, int64_t* keys_out,
                       int* ids_out, int end_bit, void* ws, size_t ws_bytes, hipStream_t s) {
    return radix_sort_pairs<uint64_t>(n, (const uint64_t*)keys_in, (const uint32_t*)ids_in,
                                      (uint64_t*)keys_out, (uint32_t*)ids_out, end_bit, ws,
                                      ws_bytes, s);
}

hipError_t map_intersects(int n, int num_isects, const float* centers, const float* extents,
                          const float* depths, const int* cum, int tiles_x, int tiles_y, int bw,
                          bool wrapped, int64_t* keys, int* ids, hipStream_t s) {
    hipError_t e = fill<int64_t>(keys, (size_t)num_isects, (int64_t)0, s);
    if (e != hipSuccess) return e;
    e = fill<int>(ids, (size_t)num_isects, 0, s);
    if (e != hipSuccess) return e;
    if (n <= 0) return hipSuccess;
    k_map_intersects<<<(n + 255) / 256, 256, 0, s>>>(n, num_isects, centers, extents, depths, cum,
                                                    tiles_x, tiles_y, bw, wrapped, keys, ids);
    return hipGetLastError();
}

hipError_t tile_bin_edges(int m, const int64_t* keys_sorted, int num_tiles, int* bins,
                          hipStream_t s) {
    hipError_t e = fill<int>(bins, (size_t)num_tiles * 2, 0, s);
    if (e != hipSuccess || m <= 0) return e;
    k_bin_edges64<<<(m + 255) / 256, 256, 0, s>>>(m, keys_sorted, num_tiles, bins);
    return hipGetLastError();
}

// ------------------------------------------------------------------ depth-first binning
// Equivalent to sorting the reference keys: a stable sort of (tile << 32 |
// depth bits) over Gaussian-major input orders each tile's entries by
// (depth bits, Gaussian id), and puts the never-written zero slots (key 0,
// id 0) first in tile 0.  So: sort the N Gaussians by depth once, emit their
// tiles in that order after the zero slots, and stable-sort by tile only.

// Per Gaussian: its depth key, and info[g] = (entries it writes (rectangle,
// capped by its count), rectangle corner x0 | y0 << 16, rectangle width, 0);
// Gaussians that write nothing sort last.  The launch also clears the tile
// bins (nbins ints) and the offset scan's look-back states (nstates words),
// and with m_exact >= 0 sets *m_total = m_exact.
__global__ __launch_bounds__(256) void k_bin_prepare(
    int n, const float* __restrict__ centers, const float* __restrict__ extents,
    const float* __restrict__ depths, const int* __restrict__ counts, int tiles_x, int tiles_y,
    int bw, uint32_t* __restrict__ depth_keys, uint32_t* __restrict__ order,
    uint4* __restrict__ info, int* __restrict__ bins, int nbins, uint64_t* __restrict__ states,
    int nstates, uint32_t* __restrict__ m_total, int m_exact) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = g; i < nbins; i += gridDim.x * blockDim.x) bins[i] = 0;
    for (int i = g; i < nstates; i += gridDim.x * blockDim.x) states[i] = 0ull;
    if (g == 0 && m_exact >= 0) *m_total = (uint32_t)m_exact;
    if (g >= n) return;
    const float2 c = reinterpret_cast<const float2*>(centers)[g];
    const float2 e = reinterpret_cast<const float2*>(extents)[g];
    int w = 0;
    uint32_t xy = 0, wd = 1;
    if (writes_intersects(e.x, e.y)) {
        int x0, y0, x1, y1;
        tile_rect(c.x, c.y, e.x, e.y, bw, tiles_x, tiles_y, x0, y0, x1, y1);
        w = min((x1 - x0) * (y1 - y0), max(counts[g], 0));
        xy = (uint32_t)x0 | ((uint32_t)y0 << 16);  // tile indices < 65536
        wd = (uint32_t)max(x1 - x0, 1);
    }
    info[g] = make_uint4((uint32_t)w, xy, wd, 0u);
    // 0xFFFFFFFF is kept for Gaussians that write nothing, so they all sort
    // after every writing one (a writing Gaussian's depth is a positive float,
    // whose bits are below 0x80000000; the clamp only touches a NaN pattern).
    depth_keys[g] = w > 0 ? min((uint32_t)__float_as_uint(depths[g]), 0xFFFFFFFEu) : 0xFFFFFFFFu;
    order[g] = (uint32_t)g;
}

// offs[p] = sum of w[order[q]] over q < p (w = info.x), and *total = the sum over
// every p, in one launch: tiles of kWrScanTile positions, 16 consecutive per
// thread, each tile adding the sum of the tiles before it by a decoupled
// look-back (ticket order, as in the one-sweep sort).  states holds one
// look-back state per tile, then the ticket and error words, all zeroed.
// rec[p] = (offs[p], order[p], its rectangle corner, its rectangle width):
// what the emission needs of position p, in one 16-byte read.

__global__ __launch_bounds__(256) void k_scan_written(int n, const uint32_t* __restrict__ order,
                                                      const uint4* __restrict__ info,
                                                      uint32_t* __restrict__ offs,
                                                      uint4* __restrict__ rec,
                                                      uint32_t* __restrict__ total,
                                                      uint64_t* __restrict__ states, int tiles) {
    __shared__ uint32_t lds[256 / kWave];
    __shared__ uint32_t s_blk, s_before;
    uint32_t* const ticket = (uint32_t*)(states + tiles);
    const int tid = threadIdx.x;
    if (tid == 0) s_blk = atomicAdd(ticket, 1u);
    __syncthreads();
    const int blk = (int)s_blk;
    const size_t base = (size_t)blk * kWrScanTile + (size_t)tid * kWrScanPer;
    uint32_t v[kWrScanPer], gs[kWrScanPer], xys[kWrScanPer], wds[kWrScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kWrScanPer; ++j) {
        gs[j] = base + j < (size_t)n ? order[base + j] : 0u;
        const uint4 f = base + j < (size_t)n ? info[gs[j]] : make_uint4(0u, 0u, 1u, 0u);
        v[j] = f.x;
        xys[j] = f.y;
        wds[j] = f.z;
        sum += v[j];
    }
    uint32_t tot;
    const uint32_t excl = block_exclusive_scan(sum, lds, tot);
    if (tid < kWave) {
        if (tid == 0)
            lb_store(states + blk, (blk == 0 ? kLbInclusive : kLbAggregate) | (uint64_t)tot);
        uint64_t before = 0;
        if (blk > 0) {
            before = lb_exclusive_wave(states, blk, ticket + 1);
            if (tid == 0) lb_store(states + blk, kLbInclusive | (before + tot));
        }
        if (tid == 0) {
            s_before = (uint32_t)before;
            if ((size_t)(blk + 1) * kWrScanTile >= (size_t)n) *total = (uint32_t)(before + tot);
        }
    }
    __syncthreads();
    uint32_t run = s_before + excl;
#pragma unroll
    for (int j = 0; j < kWrScanPer; ++j) {
        if (base + j < (size_t)n) {
            offs[base + j] = run;
            rec[base + j] = make_uint4(run, gs[j], xys[j], wds[j]);
        }
        run += v[j];
    }
}

// The (tile, id) list before the tile sort holds, by position: the zero
// slots (tile 0, id 0), which occupy [0, m - total_written), then every
// depth-ordered Gaussian's tiles.  Work on it is split by position, not by
// Gaussian: workgroup b handles positions [b * kEmitOut, (b + 1) * kEmitOut),
// so a few screen-sized splats near the camera (thousands of tiles each)
// spread over many workgroups instead of stalling one.
constexpr int kEmitOut = 2048;

// starts[b] = the depth-ordered Gaussian whose outputs hold list position
// b * out, and starts[ceil((z + total) / out)] = the last writing Gaussian,
// where z = m - total_written is the number of zero slots.  Output ranges
// [offs[p], offs[p] + w) tile [0, total) (writing Gaussians all come before
// the ones that write nothing), so every entry is written by exactly one
// lane.  Entries past max_b are not needed.
__global__ __launch_bounds__(256) void k_bin_mark_starts(int n, const uint32_t* __restrict__ offs,
                                                         const uint32_t* __restrict__ total_written,
                                                         const uint32_t* __restrict__ m_total,
                                                         int max_b, int out,
                                                         uint32_t* __restrict__ starts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t tw = *total_written, m = *m_total;
    const uint32_t z = tw <= m ? m - tw : 0u;
    const uint32_t o = z + offs[p];
    const uint32_t end = z + (p + 1 < n ? offs[p + 1] : tw);
    if (end <= o) return;
    const uint32_t uo = (uint32_t)out;
    const uint32_t b0 = (o + uo - 1) / uo, b1 = (end - 1) / uo;
    for (uint32_t b = b0; b <= b1 && b <= (uint32_t)max_b; ++b) starts[b] = (uint32_t)p;
    const uint32_t last = (z + tw + uo - 1) / uo;
    if (end == z + tw && last <= (uint32_t)max_b) starts[last] = (uint32_t)p;
}

// Total intersect count sum(max(counts, 0)) into *m_total (zeroed beforehand by
// a fill kernel): per-lane partial sums, a workgroup reduction, one atomic per
// workgroup (launched with at most kCountBlocks workgroups, so the atomics on
// the one address stay few).
constexpr int kCountBlocks = 256;
__global__ __launch_bounds__(256) void k_count_total(int n, const int* __restrict__ counts,
                                                     uint32_t* __restrict__ m_total) {
    __shared__ uint32_t wsum[256 / kWave];
    uint32_t acc = 0;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x)
        acc += (uint32_t)max(counts[g], 0);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < 256 / kWave; ++w) t += wsum[w];
        atomicAdd(m_total, t);
    }
}

// The list before the tile sort, by position in [0, cap): zero slots in
// [0, zpos), emitted output k at position z + k in [zpos, limit), sentinels
// in [limit, cap).  z = m - total_written (the zero slots a Gaussian counts
// but does not write), limit = min(m, cap).
struct ListSpan {
    uint32_t z, zpos, limit;
};
__device__ __forceinline__ ListSpan list_span(uint32_t m, uint32_t tw, uint32_t cap) {
    ListSpan sp;
    sp.z = tw <= m ? m - tw : 0u;
    sp.limit = min(m, cap);
    sp.zpos = min(sp.z, sp.limit);
    return sp;
}

// Fused emission and first tile-sort pass.  The list is never written in
// position order: k_tile_hist1 counts the first pass's digits (the low bits
// of the tile) for each block of kEmitOut positions from the Gaussians'
// rectangles, and k_emit_scatter1 generates each block's entries in LDS and
// scatters them as the first pass would.  Blocks are the k_bin_mark_starts
// blocks over positions (with m_total), so block b stages the depth-ordered
// Gaussians starts[b] .. starts[b + 1] (from 0 when the block holds zero
// slots).

// Digit counts of block b: zero slots count for tile 0, sentinels for tile
// num_tiles.  A Gaussian's outputs inside the block are a row-major stretch of
// its tile rectangle, cut into one run of consecutive tiles per rectangle row;
// a run of L tiles from tile t adds L >> bits to every digit and one to each
// of the (L mod 2^bits) digits from t mod 2^bits on, cyclically, entered in a
// difference array in LDS.  hist[d * nblocks + b] gets the count of digit d.
// Block 0 also zeroes nzero words at zero (the scan's look-back states).
__global__ __launch_bounds__(256) void k_tile_hist1(
    int n, int cap, const uint32_t* __restrict__ m_total,
    const uint32_t* __restrict__ total_written, const uint4* __restrict__ rec,
    const uint32_t* __restrict__ offs, const uint32_t* __restrict__ starts, int tiles_x,
    int num_tiles, int bits, uint32_t* __restrict__ hist, int nblocks,
    uint64_t* __restrict__ zero, int nzero) {
    __shared__ uint32_t diff[257];
    __shared__ uint32_t lds[256 / kWave];
    __shared__ uint32_t s_all;
    const int tid = threadIdx.x;
    if (blockIdx.x == 0)
        for (int i = tid; i < nzero; i += 256) zero[i] = 0ull;
    const uint32_t D = 1u << bits, dm = D - 1u;
    diff[tid] = 0u;
    if (tid == 0) {
        diff[256] = 0u;
        s_all = 0u;
    }
    __syncthreads();
    // the block's start entries are read with the totals (they are valid
    // whenever the block holds emitted entries)
    const uint32_t st0 = starts[blockIdx.x], st1 = starts[blockIdx.x + 1];
    const uint32_t tw = *total_written;
    const ListSpan sp = list_span(*m_total, tw, (uint32_t)cap);
    const uint32_t p0 = blockIdx.x * (uint32_t)kEmitOut;
    const uint32_t p1 = min(p0 + (uint32_t)kEmitOut, (uint32_t)cap);
    if (tid == 0) {
        const uint32_t ze = min(p1, sp.zpos);
        if (ze > p0) {
            atomicAdd(&diff[0], ze - p0);
            atomicAdd(&diff[1], p0 - ze);
        }
        const uint32_t s0 = max(p0, sp.limit);
        if (p1 > s0) {
            const uint32_t sd = (uint32_t)num_tiles & dm;
            atomicAdd(&diff[sd], p1 - s0);
            atomicAdd(&diff[sd + 1], s0 - p1);
        }
    }
    const uint32_t e0 = max(p0, sp.zpos), e1 = min(p1, sp.limit);
    uint32_t all = 0;
    if (e1 > e0) {
        const uint32_t k0 = e0 - sp.z, k1 = e1 - sp.z;
        const uint32_t plo = p0 <= sp.z ? 0u : st0;
        const uint32_t phi = min(st1, (uint32_t)n - 1u);
        const int cnt = phi >= plo ? (int)min(phi - plo + 1u, (uint32_t)kEmitOut + 1u) : 0;
        for (int i = tid; i < cnt; i += 256) {
            const uint32_t p = plo + (uint32_t)i;
            const uint4 r = rec[p];
            const uint32_t o = r.x;
            const uint32_t end = p + 1 < (uint32_t)n ? offs[p + 1] : tw;
            const uint32_t a = max(o, k0), bq = min(end, k1);
            if (a >= bq) continue;
            const uint32_t wd = r.w, x0 = r.z & 0xffffu, y0 = r.z >> 16;
            const uint32_t la = a - o, lb = bq - 1u - o;
            const uint32_t ra = la / wd, rb = lb / wd;
            const uint32_t ca = la - ra * wd, cb = lb - rb * wd;
            for (uint32_t row = ra; row <= rb; ++row) {
                const uint32_t cs = row == ra ? ca : 0u, ce = row == rb ? cb : wd - 1u;
                const uint32_t t0 = (y0 + row) * (uint32_t)tiles_x + x0 + cs;
                const uint32_t len = ce - cs + 1u;
                all += len >> bits;
                const uint32_t rem = len & dm;
                if (rem) {
                    const uint32_t d0 = t0 & dm, d1 = d0 + rem;
                    atomicAdd(&diff[d0], 1u);
                    if (d1 <= D) {
                        atomicAdd(&diff[d1], ~0u);
                    } else {  // wraps: digits d0 .. D - 1 and 0 .. d1 - D - 1
                        atomicAdd(&diff[0], 1u);
                        atomicAdd(&diff[d1 - D], ~0u);
                    }
                }
            }
        }
    }
    if (all) atomicAdd(&s_all, all);
    __syncthreads();
    const uint32_t v = tid < (int)D ? diff[tid] : 0u;
    uint32_t tot;
    const uint32_t incl = block_exclusive_scan(v, lds, tot) + v;
    if (tid < (int)D) hist[(size_t)tid * nblocks + blockIdx.x] = incl + s_all;
}

// Block b generates the list entries of positions [b * kEmitOut, (b + 1) *
// kEmitOut) (a zero key, the emitted key, or the sentinel key
// num_tiles << id_bits; with PACKED each emitted key is
// tile << id_bits | id, else the key is the tile and the id goes with it as
// its value), eight consecutive positions per thread, into LDS, then reads
// them back in (wave, round, lane) order and scatters them by the digit
// [id_bits, id_bits + bits) with bases from the scanned k_tile_hist1 counts.
// COMPACT stages each rectangle as one word, first tile << 12 | width (grids
// below 2^20 tiles with rows of fewer than 4096 tiles): 8 KB less LDS, one
// more workgroup per CU.
template <bool PACKED, bool COMPACT>
__global__ __launch_bounds__(256) void k_emit_scatter1(
    int n, int cap, const uint32_t* __restrict__ m_total,
    const uint32_t* __restrict__ total_written, const uint4* __restrict__ rec,
    const uint32_t* __restrict__ starts, int tiles_x, int num_tiles, int id_bits, int bits,
    const uint32_t* __restrict__ hist, int nblocks, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ ids_out) {
    // Gaussians with an output in the block: at most one per position (of
    // starts[b] .. starts[b + 1], the last one is left out when there are
    // more: it then starts after the block)
    constexpr int kStage = kEmitOut;
    constexpr int kPer = kEmitOut / 256;
    static_assert(kPer == 8 && kEmitOut == kSortThreads * 8, "one sort tile per block");
    // One LDS buffer (32 KB with packed keys and COMPACT: five workgroups per
    // CU): the staged Gaussians (offset, id, rectangle) and the generated keys
    // (and ids).  Once generated, s_off / s_g hold the digit-sorted tile and
    // the scatter's counters take the place of the rectangles.
    constexpr int kRectWords = COMPACT ? 1 : 2;
    constexpr int kWords = (2 + kRectWords) * kStage + (PACKED ? 1 : 2) * kEmitOut;
    __shared__ __align__(16) uint32_t lds[kWords];
    uint32_t* const s_off = lds;
    uint32_t* const s_g = lds + kStage;
    uint32_t* const s_r0 = lds + 2 * kStage;  // first tile << 12 | width, or corner x0 | y0 << 16
    uint32_t* const s_r1 = s_r0 + kStage;     // width (without COMPACT)
    uint32_t* const s_key = lds + (2 + kRectWords) * kStage;
    uint32_t* const s_id = s_key + kEmitOut;  // without PACKED
    static_assert(sizeof(ScatterLds) <= kRectWords * kStage * sizeof(uint32_t),
                  "counters fit the rectangles");
    ScatterLds& sl = *reinterpret_cast<ScatterLds*>(s_r0);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int ndig = 1 << bits;
    // independent loads first: this thread's digit base, the block's start
    // entries (valid whenever the block holds emitted entries) and the totals
    const uint32_t gb = tid < ndig ? hist[(size_t)tid * nblocks + blockIdx.x] : 0u;
    const uint32_t st0 = starts[blockIdx.x], st1 = starts[blockIdx.x + 1];
    const uint32_t tw = *total_written;
    const ListSpan sp = list_span(*m_total, tw, (uint32_t)cap);
    const uint32_t p0 = blockIdx.x * (uint32_t)kEmitOut;
    const int count = (int)min((uint32_t)kEmitOut, (uint32_t)cap - p0);
    const uint32_t e0 = max(p0, sp.zpos), e1 = min(p0 + (uint32_t)count, sp.limit);
    int cnt = 0;
    if (e1 > e0) {
        const uint32_t plo = p0 <= sp.z ? 0u : st0;
        const uint32_t phi = min(st1, (uint32_t)n - 1u);
        cnt = phi >= plo ? (int)min(phi - plo + 1u, (uint32_t)kStage) : 1;
        for (int i = tid; i < cnt; i += 256) {
            const uint4 r = rec[plo + (uint32_t)i];
            s_off[i] = r.x;
            s_g[i] = r.y;
            if (COMPACT) {
                const uint32_t t0 = (r.z >> 16) * (uint32_t)tiles_x + (r.z & 0xffffu);
                s_r0[i] = (t0 << 12) | r.w;
            } else {
                s_r0[i] = r.z;
                s_r1[i] = r.w;
            }
        }
    }
    __syncthreads();
    const uint32_t sentinel = (uint32_t)num_tiles << id_bits;
    const uint32_t q0 = p0 + (uint32_t)(kPer * tid);
    uint32_t ot[kPer], oi[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        ot[j] = q0 + (uint32_t)j < sp.zpos ? 0u : sentinel;
        oi[j] = 0u;
    }
    // this thread's emitted positions [a0, a1): the first one's Gaussian by a
    // binary search over the staged offsets, then a forward walk
    const uint32_t a0 = max(q0, e0), a1 = min(q0 + (uint32_t)kPer, e1);
    if (a0 < a1) {
        const uint32_t kt = a0 - sp.z;
        int lo = 0, hi = cnt - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= kt) lo = mid;
            else hi = mid - 1;
        }
        // the Gaussian's first tile t0 and rectangle width
        auto rect = [&](int at, uint32_t& t0, uint32_t& w) {
            if (COMPACT) {
                const uint32_t v = s_r0[at];
                t0 = v >> 12;
                w = v & 0xfffu;
            } else {
                const uint32_t xy = s_r0[at];
                t0 = (xy >> 16) * (uint32_t)tiles_x + (xy & 0xffffu);
                w = s_r1[at];
            }
        };
        uint32_t t0, wd, gid = s_g[lo];
        rect(lo, t0, wd);
        const uint32_t li = kt - s_off[lo];
        uint32_t ry = li / wd, rx = li - ry * wd;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t q = q0 + (uint32_t)j;
            if (q >= a0 && q < a1) {
                const uint32_t k = q - sp.z;
                if (lo + 1 < cnt && s_off[lo + 1] <= k) {
                    while (lo + 1 < cnt && s_off[lo + 1] <= k) ++lo;
                    rect(lo, t0, wd);
                    gid = s_g[lo];
                    ry = 0;
                    rx = 0;
                }
                const uint32_t tile = t0 + ry * (uint32_t)tiles_x + rx;
                if (PACKED) {
                    ot[j] = (tile << id_bits) | gid;
                } else {
                    ot[j] = tile;
                    oi[j] = gid;
                }
                if (++rx == wd) {
                    rx = 0;
                    ++ry;
                }
            }
        }
    }
    {
        uint4* const dk = reinterpret_cast<uint4*>(s_key + kPer * tid);
        dk[0] = make_uint4(ot[0], ot[1], ot[2], ot[3]);
        dk[1] = make_uint4(ot[4], ot[5], ot[6], ot[7]);
        if (!PACKED) {
            uint4* const di = reinterpret_cast<uint4*>(s_id + kPer * tid);
            di[0] = make_uint4(oi[0], oi[1], oi[2], oi[3]);
            di[1] = make_uint4(oi[4], oi[5], oi[6], oi[7]);
        }
    }
    __syncthreads();
    uint32_t key[kPer], val[kPer];
    const int wbase = wave * kPer * kWave;
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int li = wbase + r * kWave + lane;
        key[r] = s_key[li];
        val[r] = PACKED ? 0u : s_id[li];
    }
    tile_rank_scatter<uint32_t, kPer>(key, val, count, !PACKED, 0u, id_bits, bits, gb, keys_out,
                                      ids_out, (uint32_t)cap, sl, s_off, s_g);
}

static inline int bits_for(int v) {  // bits needed to represent 0..v-1
    int b = 0;
    while ((1ll << b) < (long long)v) ++b;
    return b;
}

static int emit_blocks(int cap) { return (cap + kEmitOut - 1) / kEmitOut; }

// first tile-pass digit counts: up to 256 digits per emission block, then the
// look-back states of their scan
static size_t hist1_words(int cap) { return (size_t)256 * emit_blocks(cap); }

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Binning runs in two phases.  The depth phase needs only the Gaussians:
// entry counts and rectangles, the depth sort and the emission offsets and
// records.  The tile phase needs the list size as well: the fused emission
// and first tile pass, the remaining passes and the bin edges.  Each has its
// own workspace (the tile phase reads the depth phase's), so a caller that
// reads the intersect total back to the host can queue the depth phase first.
struct DepthWs {
    uint32_t *dkey, *dkey2, *ord, *ord2;
    uint4 *info, *rec;
    uint32_t* offs;     // n offsets, then the written total and the intersect total
    uint64_t* states;   // offset scan look-back states
    void* rws;
    size_t rws_bytes, total;
};

static DepthWs depth_ws(int n, void* base) {
    char* b = (char*)base;
    DepthWs d;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        char* p = b ? b + at : nullptr;
        at += up256(bytes);
        return p;
    };
    d.dkey = (uint32_t*)take(4 * (size_t)n);
    d.dkey2 = (uint32_t*)take(4 * (size_t)n);
    d.ord = (uint32_t*)take(4 * (size_t)n);
    d.ord2 = (uint32_t*)take(4 * (size_t)n);
    d.info = (uint4*)take(16 * (size_t)n);
    d.rec = (uint4*)take(16 * (size_t)n);
    d.offs = (uint32_t*)take(4 * (size_t)n + 8);
    d.states = (uint64_t*)take(8 * ((size_t)wr_scan_tiles(n) + 1));
    d.rws_bytes = up256(radix_workspace_bytes<uint32_t>(n));
    d.rws = take(d.rws_bytes);
    d.total = at;
    return d;
}

struct TileWs {
    uint32_t *tkey, *tkey2, *tid_in, *starts, *hist1, *m_word;
    uint64_t* hstates;
    void* rws;
    size_t rws_bytes, total;
};

static TileWs tile_ws(int cap, void* base) {
    char* b = (char*)base;
    TileWs t;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        char* p = b ? b + at : nullptr;
        at += up256(bytes);
        return p;
    };
    t.tkey = (uint32_t*)take(4 * (size_t)cap);
    t.tkey2 = (uint32_t*)take(4 * (size_t)cap);
    t.tid_in = (uint32_t*)take(4 * (size_t)cap);  // ids after the first tile pass
    t.starts = (uint32_t*)take(4 * ((size_t)emit_blocks(cap) + 2));
    t.hist1 = (uint32_t*)take(4 * hist1_words(cap));
    t.hstates = (uint64_t*)take(8 * ((size_t)wr_scan_tiles((int)hist1_words(cap)) + 1));
    t.m_word = (uint32_t*)take(4);
    t.rws_bytes = up256(radix_workspace_bytes<uint32_t>(cap));
    t.rws = take(t.rws_bytes);
    t.total = at;
    return t;
}

size_t bin_depth_workspace_bytes(int n) { return depth_ws(std::max(n, 1), nullptr).total; }
size_t bin_tiles_workspace_bytes(int cap) { return tile_ws(std::max(cap, 1), nullptr).total; }
size_t bin_gaussians_workspace_bytes(int n, int m) {
    return bin_depth_workspace_bytes(n) + bin_tiles_workspace_bytes(m);
}

hipError_t count_intersects(int n, const int* counts, unsigned* out, hipStream_t s) {
    hipError_t e = fill<uint32_t>(out, 1, 0u, s);
    if (e != hipSuccess || n <= 0) return e;
    k_count_total<<<std::min((n + 255) / 256, kCountBlocks), 256, 0, s>>>(n, counts, out);
    return hipGetLastError();
}

// Depth phase (n > 0).  Also clears the tile bins; with m_exact >= 0 the
// intersect total word of the workspace is set to it.
static hipError_t bin_depth_phase(int n, const float* centers, const float* extents,
                                  const float* depths, const int* counts, int tiles_x,
                                  int tiles_y, int bw, int* tile_bins, const DepthWs& d,
                                  int m_exact, hipStream_t s) {
    const int gn = (n + 255) / 256, st = wr_scan_tiles(n);
    uint32_t* total = d.offs + n;  // entries written by Gaussians
    k_bin_prepare<<<gn, 256, 0, s>>>(n, centers, extents, depths, counts, tiles_x, tiles_y, bw,
                                     d.dkey, d.ord, d.info, tile_bins, 2 * tiles_x * tiles_y,
                                     d.states, st + 1, d.offs + n + 1, m_exact);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs<uint32_t>(n, d.dkey, d.ord, d.dkey2, d.ord2, 32, d.rws, d.rws_bytes, s);
    if (e != hipSuccess) return e;
    k_scan_written<<<st, 256, 0, s>>>(n, d.ord2, d.info, d.offs, d.rec, total, d.states, st);
    return hipGetLastError();
}

// Tile phase (n > 0, cap > 0), with the intersect total (list entries, zero
// slots included) read from m_total on the device.
static hipError_t bin_tiles_phase(int n, int cap, const uint32_t* m_total, int tiles_x,
                                  int tiles_y, const DepthWs& d, int* ids_sorted, int* tile_bins,
                                  const TileWs& t, hipStream_t s) {
    const int num_tiles = tiles_x * tiles_y, gn = (n + 255) / 256;
    const uint32_t* total = d.offs + n;
    // the sentinel tile (num_tiles) needs one more key value; when the tile
    // and the Gaussian id fit one 32-bit word together, the entries are sorted
    // as packed keys alone (half the bytes per pass), and the ids come out of
    // the bin-edge pass
    const int tile_bits = bits_for(num_tiles + 1), id_bits = bits_for(n);
    const int pk = (id_bits > 0 && tile_bits + id_bits <= 32) ? id_bits : 0;
    // the emission is fused with the first tile pass (digit width bits1, as
    // the sort would split tile_bits); later passes run as a sort from bit
    // pk + bits1
    const int tpasses = (tile_bits + 7) / 8;
    const int bits1 = (tile_bits + tpasses - 1) / tpasses;
    const int eb = emit_blocks(cap);
    const int hcount = (1 << bits1) * eb;
    const bool hlb = hcount > kScanSingleMax;
    const int hlbt = wr_scan_tiles(hcount);
    hipError_t e;
    k_bin_mark_starts<<<gn, 256, 0, s>>>(n, d.offs, total, m_total, eb, kEmitOut, t.starts);
    k_tile_hist1<<<eb, 256, 0, s>>>(n, cap, m_total, total, d.rec, d.offs, t.starts, tiles_x,
                                    num_tiles, bits1, t.hist1, eb, hlb ? t.hstates : nullptr,
                                    hlb ? hlbt + 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (hlb) {
        k_scan_lookback<<<hlbt, 256, 0, s>>>(t.hist1, hcount, t.hstates, hlbt);
    } else if ((e = scan_exclusive(t.hist1, hcount, nullptr, nullptr, s)) != hipSuccess) {
        return e;
    }
    const bool more = tile_bits > bits1;
    uint32_t* k1 = more ? t.tkey : t.tkey2;
    // rectangles staged as one word when the tile grid allows it
    const bool compact = num_tiles < (1 << 20) && tiles_x < 4096;
    uint32_t* const v1 = pk ? nullptr : (more ? t.tid_in : (uint32_t*)ids_sorted);
#define GSTEX_EMIT(P, C)                                                                    \
    k_emit_scatter1<P, C><<<eb, 256, 0, s>>>(n, cap, m_total, total, d.rec, t.starts, tiles_x, \
                                             num_tiles, pk, bits1, t.hist1, eb, k1, v1)
    if (pk && compact) GSTEX_EMIT(true, true);
    else if (pk) GSTEX_EMIT(true, false);
    else if (compact) GSTEX_EMIT(false, true);
    else GSTEX_EMIT(false, false);
#undef GSTEX_EMIT
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int bits2 = tile_bits - bits1;
    if (more && pk && bits2 <= 8) {
        // one pass left: it writes the ids and the bin edges itself
        // (k_last_pass_edges, k_run_edges), with its digit counts and their
        // scan in the tile sort's workspace
        const int nb2 = sort_blocks(cap), shift2 = pk + bits1;
        const bool small = sort_items(cap) == 8;
        uint32_t* const hist2 = (uint32_t*)t.rws;
        uint64_t* const st2 = (uint64_t*)((char*)t.rws + up256(4 * 256 * (size_t)nb2));
        const int hc2 = (1 << bits2) * nb2;
        const bool lb2 = hc2 > kScanSingleMax;
        const int lbt2 = wr_scan_tiles(hc2);
        if (small)
            k_radix_hist<uint32_t, 8><<<nb2, kSortThreads, 0, s>>>(
                cap, t.tkey, 0u, shift2, bits2, hist2, nb2, lb2 ? st2 : nullptr, lb2 ? lbt2 + 1 : 0);
        else
            k_radix_hist<uint32_t, 16><<<nb2, kSortThreads, 0, s>>>(
                cap, t.tkey, 0u, shift2, bits2, hist2, nb2, lb2 ? st2 : nullptr, lb2 ? lbt2 + 1 : 0);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (lb2) {
            k_scan_lookback<<<lbt2, 256, 0, s>>>(hist2, hc2, st2, lbt2);
        } else if ((e = scan_exclusive(hist2, hc2, nullptr, nullptr, s)) != hipSuccess) {
            return e;
        }
        if (small)
            k_last_pass_edges<8><<<nb2, kSortThreads, 0, s>>>(cap, t.tkey, t.tkey2, ids_sorted, pk,
                                                              shift2, bits2, hist2, nb2, num_tiles,
                                                              tile_bins);
        else
            k_last_pass_edges<16><<<nb2, kSortThreads, 0, s>>>(cap, t.tkey, t.tkey2, ids_sorted,
                                                               pk, shift2, bits2, hist2, nb2,
                                                               num_tiles, tile_bins);
        k_run_edges<<<(hc2 + 255) / 256, 256, 0, s>>>(cap, hc2, hist2, t.tkey2, pk, num_tiles,
                                                      tile_bins);
        return hipGetLastError();
    }
    if (more) {
        if (pk)
            e = radix_sort_pairs<uint32_t>(cap, t.tkey, nullptr, t.tkey2, nullptr,
                                           pk + tile_bits, t.rws, t.rws_bytes, s, pk + bits1);
        else
            e = radix_sort_pairs<uint32_t>(cap, t.tkey, t.tid_in, t.tkey2, (uint32_t*)ids_sorted,
                                           tile_bits, t.rws, t.rws_bytes, s, bits1);
        if (e != hipSuccess) return e;
    }
    k_bin_edges32<<<(cap + 1023) / 1024, 256, 0, s>>>(
        cap, t.tkey2, num_tiles, tile_bins,
        (reinterpret_cast<uintptr_t>(t.tkey2) & 15) == 0 &&
            (reinterpret_cast<uintptr_t>(ids_sorted) & 15) == 0,
        pk, ids_sorted);
    return hipGetLastError();
}

hipError_t bin_depth(int n, const float* centers, const float* extents, const float* depths,
                     const int* counts, int tiles_x, int tiles_y, int bw, int* tile_bins,
                     void* ws, size_t ws_bytes, hipStream_t s) {
    if (n <= 0) return fill<int>(tile_bins, (size_t)tiles_x * tiles_y * 2, 0, s);
    const DepthWs d = depth_ws(n, ws);
    if (ws_bytes < d.total) return hipErrorInvalidValue;
    return bin_depth_phase(n, centers, extents, depths, counts, tiles_x, tiles_y, bw, tile_bins, d,
                           -1, s);
}

hipError_t bin_tiles(int n, int cap, const unsigned* m_total, int tiles_x, int tiles_y,
                     const void* depth_ws_base, size_t depth_ws_bytes, int* ids_sorted,
                     int* tile_bins, void* ws, size_t ws_bytes, hipStream_t s) {
    if (n <= 0 || cap <= 0) return hipSuccess;  // the depth phase left the bins empty
    const DepthWs d = depth_ws(n, const_cast<void*>(depth_ws_base));
    const TileWs t = tile_ws(cap, ws);
    if (depth_ws_bytes < d.total || ws_bytes < t.total) return hipErrorInvalidValue;
    const uint32_t* mt = m_total;
    if (!mt) {
        hipError_t e = fill<uint32_t>(t.m_word, 1, (uint32_t)cap, s);
        if (e != hipSuccess) return e;
        mt = t.m_word;
    }
    return bin_tiles_phase(n, cap, mt, tiles_x, tiles_y, d, ids_sorted, tile_bins, t, s);
}

hipError_t bin_gaussians(int n, int cap, const float* centers, const float* extents,
                         const float* depths, const int* counts, int tiles_x, int tiles_y, int bw,
                         int* ids_sorted, int* tile_bins, unsigned* m_out, void* ws,
                         size_t ws_bytes, hipStream_t s) {
    const int num_tiles = tiles_x * tiles_y;
    hipError_t e = hipSuccess;
    if (n <= 0 || cap <= 0) e = fill<int>(tile_bins, (size_t)num_tiles * 2, 0, s);
    if (e != hipSuccess || n <= 0) return e;
    if (cap <= 0) {
        if (m_out) {
            // the total is still reported, so the caller can size a capacity
            if (ws_bytes < bin_gaussians_workspace_bytes(n, 1)) return hipErrorInvalidValue;
            uint32_t* mt = depth_ws(n, ws).offs + n + 1;
            if ((e = fill<uint32_t>(mt, 1, 0u, s)) != hipSuccess) return e;
            k_count_total<<<std::min((n + 255) / 256, kCountBlocks), 256, 0, s>>>(n, counts, mt);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            return hipMemcpyAsync(m_out, mt, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        }
        return hipSuccess;
    }
    if (ws_bytes < bin_gaussians_workspace_bytes(n, cap)) return hipErrorInvalidValue;
    const DepthWs d = depth_ws(n, ws);
    const TileWs t = tile_ws(cap, (char*)ws + bin_depth_workspace_bytes(n));
    uint32_t* m_total = d.offs + n + 1;  // intersect total (written + zero slots)
    // capacity mode: the total is counted on the device; exact size: the
    // caller's num_intersects is that total (a wrong one still keeps every
    // write inside the list and the zero-slot range), set by k_bin_prepare
    if (m_out) {
        if ((e = fill<uint32_t>(m_total, 1, 0u, s)) != hipSuccess) return e;
        k_count_total<<<std::min((n + 255) / 256, kCountBlocks), 256, 0, s>>>(n, counts, m_total);
    }
    e = bin_depth_phase(n, centers, extents, depths, counts, tiles_x, tiles_y, bw, tile_bins, d,
                        m_out ? -1 : cap, s);
    if (e != hipSuccess) return e;
    e = bin_tiles_phase(n, cap, m_total, tiles_x, tiles_y, d, ids_sorted, tile_bins, t, s);
    if (e != hipSuccess) return e;
    if (m_out) return hipMemcpyAsync(m_out, m_total, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    return hipSuccess;
}

}  // namespace gstex
