// ft_build.hip -- gfx950 kernels of the HLBVH build (build_bvh, src/hlbvh.hpp:2165-2302):
//   bbox -> SFC encode -> stable LSD radix sort (key u64 + index u32) -> tree assembly
//   (leaf gather, topology and boxes in one bottom-up pass per chunk of leaves, then
//   the nodes that span chunks).
// Every stage is stream-ordered; nothing synchronises with the host.
#include "ft_internal.hpp"

#include <algorithm>
#include <vector>

namespace ft {

// ===========================================================================
// Peano-Hilbert state machine (the curve of src/sfc.peano_hilbert.hpp:332-376):
// entry [state*8 + pix] = next_state << 3 | 3-bit digit.
// ===========================================================================
__constant__ uint16_t c_ph_table[384] = {
    0x120, 0x0e7, 0x0c9, 0x0de, 0x053, 0x054, 0x0ca, 0x0dd, 0x0ef, 0x05c, 0x0c6, 0x0c5, 0x128, 0x05b, 0x0d1, 0x0d2, 0x044, 0x043, 0x0cd, 0x0da, 0x0f7, 0x130, 0x0ce, 0x0d9, 0x04b, 0x138, 0x0c2, 0x0c1, 0x04c, 0x0ff, 0x0d5, 0x0d6,
    0x141, 0x0c0, 0x166, 0x107, 0x142, 0x033, 0x165, 0x034, 0x0c8, 0x03b, 0x10f, 0x03c, 0x149, 0x14a, 0x16e, 0x16d, 0x023, 0x152, 0x024, 0x175, 0x0d0, 0x151, 0x117, 0x176, 0x15a, 0x159, 0x17d, 0x17e, 0x02b, 0x0d8, 0x02c, 0x11f,
    0x10e, 0x119, 0x127, 0x0e0, 0x10d, 0x11a, 0x014, 0x013, 0x101, 0x102, 0x0e8, 0x01b, 0x116, 0x115, 0x12f, 0x01c, 0x10a, 0x11d, 0x003, 0x004, 0x109, 0x11e, 0x0f0, 0x137, 0x105, 0x106, 0x00c, 0x13f, 0x112, 0x111, 0x00b, 0x0f8,
    0x0c7, 0x156, 0x100, 0x171, 0x074, 0x155, 0x073, 0x172, 0x15e, 0x15d, 0x179, 0x17a, 0x0cf, 0x07c, 0x108, 0x07b, 0x145, 0x064, 0x162, 0x063, 0x146, 0x0d7, 0x161, 0x110, 0x06c, 0x0df, 0x06b, 0x118, 0x14d, 0x14e, 0x16a, 0x169,
    0x0e6, 0x14f, 0x0e5, 0x0b4, 0x131, 0x158, 0x132, 0x0b3, 0x157, 0x140, 0x0bc, 0x0bb, 0x0ee, 0x139, 0x0ed, 0x13a, 0x148, 0x121, 0x0a3, 0x122, 0x15f, 0x0f6, 0x0a4, 0x0f5, 0x129, 0x0fe, 0x12a, 0x0fd, 0x150, 0x147, 0x0ab, 0x0ac,
    0x0e2, 0x093, 0x0e1, 0x168, 0x135, 0x094, 0x136, 0x17f, 0x09b, 0x09c, 0x170, 0x167, 0x0ea, 0x13d, 0x0e9, 0x13e, 0x084, 0x125, 0x16f, 0x126, 0x083, 0x0f2, 0x178, 0x0f1, 0x12d, 0x0fa, 0x12e, 0x0f9, 0x08c, 0x08b, 0x177, 0x160,
    0x067, 0x020, 0x00e, 0x019, 0x114, 0x113, 0x00d, 0x01a, 0x028, 0x11b, 0x001, 0x002, 0x06f, 0x11c, 0x016, 0x015, 0x103, 0x104, 0x00a, 0x01d, 0x030, 0x077, 0x009, 0x01e, 0x10c, 0x07f, 0x005, 0x006, 0x10b, 0x038, 0x012, 0x011,
    0x086, 0x007, 0x0a1, 0x040, 0x085, 0x0f4, 0x0a2, 0x0f3, 0x00f, 0x0fc, 0x048, 0x0fb, 0x08e, 0x08d, 0x0a9, 0x0aa, 0x0e4, 0x095, 0x0e3, 0x0b2, 0x017, 0x096, 0x050, 0x0b1, 0x09d, 0x09e, 0x0ba, 0x0b9, 0x0ec, 0x01f, 0x0eb, 0x058,
    0x049, 0x05e, 0x060, 0x027, 0x04a, 0x05d, 0x0d3, 0x0d4, 0x046, 0x045, 0x02f, 0x0dc, 0x051, 0x052, 0x068, 0x0db, 0x04d, 0x05a, 0x0c4, 0x0c3, 0x04e, 0x059, 0x037, 0x070, 0x042, 0x041, 0x0cb, 0x078, 0x055, 0x056, 0x0cc, 0x03f,
    0x000, 0x091, 0x047, 0x0b6, 0x133, 0x092, 0x134, 0x0b5, 0x099, 0x09a, 0x0be, 0x0bd, 0x008, 0x13b, 0x04f, 0x13c, 0x082, 0x123, 0x0a5, 0x124, 0x081, 0x010, 0x0a6, 0x057, 0x12b, 0x018, 0x12c, 0x05f, 0x08a, 0x089, 0x0ad, 0x0ae,
    0x021, 0x088, 0x022, 0x173, 0x076, 0x09f, 0x075, 0x174, 0x090, 0x087, 0x17b, 0x17c, 0x029, 0x07e, 0x02a, 0x07d, 0x08f, 0x066, 0x164, 0x065, 0x098, 0x031, 0x163, 0x032, 0x06e, 0x039, 0x06d, 0x03a, 0x097, 0x080, 0x16c, 0x16b,
    0x025, 0x154, 0x026, 0x0af, 0x072, 0x153, 0x071, 0x0b8, 0x15c, 0x15b, 0x0b7, 0x0a0, 0x02d, 0x07a, 0x02e, 0x079, 0x143, 0x062, 0x0a8, 0x061, 0x144, 0x035, 0x0bf, 0x036, 0x06a, 0x03d, 0x069, 0x03e, 0x14b, 0x14c, 0x0b0, 0x0a7,
};

// Morton bit spread: bit b of a 21-bit value moves to bit 3b (spread3_u64,
// src/sfc.morton.hpp:33-41), by the usual shift-and-mask cascade.
__device__ __host__ inline uint64_t spread3_21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x001F00000000FFFFull;
    v = (v | (v << 16)) & 0x001F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

template <typename T> struct BboxDev {
    T lo[3];
    T hi[3];
    T inv[3];
};

// ===========================================================================
// K1: bounding box (compute_bbox, src/hlbvh.hpp:462-578): exact min / max.
// ===========================================================================
template <typename T> __device__ inline T tmin(T a, T b) { return b < a ? b : a; }
template <typename T> __device__ inline T tmax(T a, T b) { return a < b ? b : a; }

template <typename T> __device__ inline T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = tmin(v, __shfl_xor(v, o, 64));
    return v;
}
template <typename T> __device__ inline T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = tmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <typename T>
__global__ __launch_bounds__(256) void bbox_partial_kernel(const T *__restrict__ x,
                                                           const T *__restrict__ y,
                                                           const T *__restrict__ z, size_t n,
                                                           T *__restrict__ partials) {
    T v[6];
    {
        T f = x[0], g = y[0], h = z[0];
        v[0] = f; v[1] = f; v[2] = g; v[3] = g; v[4] = h; v[5] = h;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        T a = x[i], b = y[i], c = z[i];
        v[0] = tmin(v[0], a); v[1] = tmax(v[1], a);
        v[2] = tmin(v[2], b); v[3] = tmax(v[3], b);
        v[4] = tmin(v[4], c); v[5] = tmax(v[5], c);
    }
    __shared__ T red[4][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        T r = (a & 1) ? wave_max(v[a]) : wave_min(v[a]);
        if (lane == 0) red[w][a] = r;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        T r = red[0][a];
        for (int q = 1; q < 4; ++q) r = (a & 1) ? tmax(r, red[q][a]) : tmin(r, red[q][a]);
        partials[blockIdx.x * 6 + a] = r;
    }
}

// Reduces the partials (or takes a caller-supplied box) and stores lo/hi/inv.
// inv = 0 for a flat axis, else 1/extent in T (src/hlbvh.hpp:392-394).
template <typename T>
__global__ __launch_bounds__(64) void bbox_final_kernel(const T *__restrict__ partials, int np,
                                                        int use_given, T g0, T g1, T g2, T g3,
                                                        T g4, T g5, BboxDev<T> *out) {
    if (threadIdx.x != 0) return;
    T v[6];
    if (use_given) {
        v[0] = g0; v[1] = g1; v[2] = g2; v[3] = g3; v[4] = g4; v[5] = g5;
    } else {
        for (int a = 0; a < 6; ++a) v[a] = partials[a];
        for (int b = 1; b < np; ++b)
            for (int a = 0; a < 6; ++a)
                v[a] = (a & 1) ? tmax(v[a], partials[b * 6 + a]) : tmin(v[a], partials[b * 6 + a]);
    }
    for (int a = 0; a < 3; ++a) {
        out->lo[a] = v[2 * a];
        out->hi[a] = v[2 * a + 1];
        if constexpr (CoordTraits<T>::fp) {
            T d = v[2 * a + 1] - v[2 * a];
            out->inv[a] = (d == T(0)) ? T(0) : (T(1) / d);
        } else {
            out->inv[a] = T(0);
        }
    }
}

// ===========================================================================
// K3/K4: SFC encode (sfc_encode, src/hlbvh.hpp:357-423).  One particle per lane,
// coalesced SoA reads, PH state table staged in LDS.
// ===========================================================================
template <typename T> __device__ inline uint32_t quantise(T nx, int B, int sfc);

// Morton FP: top B mantissa bits of fl(1 + nx) in T (encode_to_sfc1d,
// src/hlbvh.hpp:246-271); PH: of min(1 + (double)nx, 0x1.fffffffffffffp+0)
// (quantize_coord, src/sfc.peano_hilbert.hpp:472-476).
template <> __device__ inline uint32_t quantise<double>(double nx, int B, int sfc) {
    if (sfc == FT_SFC_MORTON) {
        double v = 1.0 + nx;
        uint64_t bits = (uint64_t)__double_as_longlong(v);
        return (uint32_t)((bits & 0x000FFFFFFFFFFFFFull) >> (52 - B));
    }
    double v = fmin(1.0 + nx, 0x1.fffffffffffffp+0);
    uint64_t bits = (uint64_t)__double_as_longlong(v);
    return (uint32_t)(((bits & 0x000FFFFFFFFFFFFFull) >> (52 - B)) & ((1ull << B) - 1));
}
template <> __device__ inline uint32_t quantise<float>(float nx, int B, int sfc) {
    if (sfc == FT_SFC_MORTON) {
        float v = 1.0f + nx;
        uint32_t bits = (uint32_t)__float_as_uint(v);
        return (bits & 0x007FFFFFu) >> (23 - B);
    }
    double v = fmin(1.0 + (double)nx, 0x1.fffffffffffffp+0);
    uint64_t bits = (uint64_t)__double_as_longlong(v);
    return (uint32_t)(((bits & 0x000FFFFFFFFFFFFFull) >> (52 - B)) & ((1ull << B) - 1));
}

__device__ inline uint64_t ph_encode(uint32_t ix, uint32_t iy, uint32_t iz, int B,
                                     const uint16_t *tab) {
    uint64_t key = 0;
    uint32_t state = 0;
    for (int bit = B - 1; bit >= 0; --bit) {
        uint32_t pix = (((ix >> bit) & 1u) << 2) | (((iy >> bit) & 1u) << 1) | ((iz >> bit) & 1u);
        uint32_t e = tab[state * 8 + pix];
        key = (key << 3) | (e & 7u);
        state = e >> 3;
    }
    return key;
}

// A point as one aligned record of four coordinates (x, y, z, 0).
template <typename T> struct alignas(4 * sizeof(T)) Point4 {
    T v[4];
};

// Also accumulates the radix-sort digit histograms of every pass (digits of
// rbits bits; pass p's bins at hist[p << rbits ...]) so the sort does not read
// the keys once more just to count them.  With `pts`, it also writes each
// point as one record for the leaf gather.
constexpr int ENC_HIST_WORDS = 8 * 512;  // up to 8 passes of 9-bit digits

// The SFC key of one point (box lo / inv: FP coordinates only).
template <typename T>
__device__ __forceinline__ uint64_t encode_point(const T p[3], int B, int sfc, const T lo[3],
                                                 const T inv[3], const uint16_t *tab) {
    const T clamp_hi = CoordTraits<T>::fp ? (T)(sizeof(T) == 8 ? 0x1.fffffffffffffp-1 : 0x1.fffffep-1) : T(0);
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if constexpr (CoordTraits<T>::fp) {
            T t = (p[a] - lo[a]) * inv[a];
            T nx = ffmin(ffmax(t, T(0)), clamp_hi);  // sycl::clamp
            q[a] = quantise<T>(nx, B, sfc);
        } else {
            q[a] = (uint32_t)((uint64_t)p[a] & ((1ull << B) - 1));  // x & (2^B-1)
        }
    }
    if (sfc == FT_SFC_MORTON) return spread3_21(q[0]) | (spread3_21(q[1]) << 1) | (spread3_21(q[2]) << 2);
    return ph_encode(q[0], q[1], q[2], B, tab);
}

// Adds one key's digits of every pass to the block's LDS histograms.  When a
// wave's active keys share a digit (the high digits of spatially clustered
// input runs), one add of the lane count replaces that many atomics on the
// same LDS word.
__device__ __forceinline__ void hist_add_key(uint32_t *h, uint64_t key, int passes, int rbits) {
    const uint32_t dmask = (1u << rbits) - 1u;
    const uint64_t act = __ballot(1);
    const int lead = __ffsll((long long)act) - 1;
    for (int p = 0; p < passes; ++p) {
        const uint32_t dg = (uint32_t)(key >> (rbits * p)) & dmask;
        const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)dg);
        if (__ballot(dg == d0) == act) {
            if ((int)(threadIdx.x & 63) == lead)
                atomicAdd(&h[(p << rbits) + d0], (uint32_t)__popcll(act));
        } else {
            atomicAdd(&h[(p << rbits) + dg], 1u);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void encode_kernel(const T *__restrict__ x,
                                                     const T *__restrict__ y,
                                                     const T *__restrict__ z, size_t n, int B,
                                                     int sfc, const BboxDev<T> *__restrict__ bb,
                                                     uint64_t *__restrict__ keys, int passes,
                                                     int rbits, uint32_t *__restrict__ hist,
                                                     Point4<T> *__restrict__ pts) {
    __shared__ uint16_t tab[384];
    __shared__ uint32_t h[ENC_HIST_WORDS];
    for (int i = threadIdx.x; i < 384; i += blockDim.x) tab[i] = c_ph_table[i];
    for (int i = threadIdx.x; i < ENC_HIST_WORDS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    T lo[3] = {T(0), T(0), T(0)}, inv[3] = {T(0), T(0), T(0)};
    if constexpr (CoordTraits<T>::fp) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = bb->lo[a];
            inv[a] = bb->inv[a];
        }
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const T p[3] = {x[i], y[i], z[i]};
        if (pts) pts[i] = Point4<T>{{p[0], p[1], p[2], T(0)}};
        const uint64_t key = encode_point<T>(p, B, sfc, lo, inv, tab);
        keys[i] = key;
        if (hist) hist_add_key(h, key, passes, rbits);
    }
    if (hist) {
        __syncthreads();
        for (int i = threadIdx.x; i < (passes << rbits); i += blockDim.x) {
            const uint32_t v = h[i];
            if (v) atomicAdd(&hist[i], v);
        }
    }
}

// ===========================================================================
// K7: stable LSD radix sort of (key u64, index u32), digits of BITS = 8 or 9
// bits, one sweep per digit with decoupled look-back (tile prefixes published
// per digit, read back by later tiles), wave-level multisplit ranking (BITS
// ballots per item) and an LDS-staged scatter so each digit's run is written
// contiguously.  Equal keys keep their input order (stable), so every digit
// width gives the same permutation.
// ===========================================================================
constexpr int RS_BLOCK = 512;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ITEMS = 12;                   // keys per thread
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;   // 6144 keys per tile
constexpr int RS_TILE_MIN = 2048;              // smallest tile: sizes the look-back words
static_assert(RS_TILE_MIN <= RS_TILE, "look-back words are sized for tiles of at least RS_TILE_MIN keys");
constexpr int RS_MAX_PASSES = 8;
constexpr int RS_MAX_RADIX = 512;              // 9-bit digits
constexpr int RS_HIST_WORDS = RS_MAX_PASSES * RS_MAX_RADIX;
static_assert(RS_MAX_RADIX <= RS_BLOCK, "one thread per digit");
static_assert(RS_HIST_WORDS <= ENC_HIST_WORDS, "encode histograms");

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Global digit histograms of every pass in one read of the keys: pass p's
// bins are hist[p << rbits ...].
__global__ __launch_bounds__(256) void radix_hist_kernel(const uint64_t *__restrict__ keys,
                                                         size_t n, int passes, int rbits,
                                                         uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[RS_HIST_WORDS];
    const uint32_t mask = (1u << rbits) - 1u;
    for (int i = threadIdx.x; i < RS_HIST_WORDS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t k = keys[i];
        for (int p = 0; p < passes; ++p)
            atomicAdd(&h[(p << rbits) + ((uint32_t)(k >> (rbits * p)) & mask)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (passes << rbits); i += blockDim.x) {
        uint32_t v = h[i];
        if (v) atomicAdd(&hist[i], v);
    }
}

// Exclusive scan over the block (any number of waves): s_tmp holds one word
// per wave.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *s_tmp) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int q = 0; q < w; ++q) add += s_tmp[q];
    __syncthreads();
    return x - v + add;
}

// Exclusive scan of each pass's histogram -> digit base offsets (one thread
// per digit: launched with 1 << rbits threads); resets the per-pass tile
// counters and advances the sort counter (epoch) that tags this sort's
// look-back words.  Epochs run 1 .. 2^27-1 and then wrap to 1.
__global__ __launch_bounds__(RS_MAX_RADIX) void radix_base_kernel(
    const uint32_t *__restrict__ hist, int passes, int rbits, uint32_t *__restrict__ base,
    uint32_t *__restrict__ tile_ctr, uint32_t *__restrict__ epoch) {
    __shared__ uint32_t tmp[RS_MAX_RADIX / 64];
    for (int p = 0; p < passes; ++p) {
        uint32_t v = hist[(p << rbits) + threadIdx.x];
        base[(p << rbits) + threadIdx.x] = block_excl_scan(v, tmp);
    }
    if (threadIdx.x < RS_MAX_PASSES) tile_ctr[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        const uint32_t e = *epoch + 1u;
        *epoch = (e >= (1u << 27)) ? 1u : e;
    }
}

__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// SEG: many trees' keys at once.  seg_tiles[tile] = (first key, end, tree,
// first tile of its tree); tiles never cross trees, a tree's digit bases
// (which include its first index) are digit_base + tree * RS_HIST_WORDS, and
// the look-back stops at the tree's first tile.
template <bool FIRST, int BITS, bool SEG = false>
__global__ __launch_bounds__(RS_BLOCK) void onesweep_kernel(
    const uint64_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint64_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, int n, int shift,
    const uint32_t *__restrict__ digit_base, uint64_t *__restrict__ lookback,
    const uint32_t *__restrict__ epoch, int pass, uint32_t *__restrict__ tile_ctr,
    const int4 *__restrict__ seg_tiles = nullptr) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t MASK = RADIX - 1;
    static_assert(RADIX <= RS_MAX_RADIX, "digit width");
    using dig_t = typename std::conditional<(BITS <= 8), uint8_t, uint16_t>::type;
    // One staging buffer: the per-wave digit counters while the tile is
    // ranked, then the keys, then the values (as 32-bit words); each
    // position's digit is kept for the values.
    constexpr int STAGE_BYTES = RS_TILE * 8 > RS_WAVES * RADIX * 4 ? RS_TILE * 8 : RS_WAVES * RADIX * 4;
    __shared__ __align__(16) uint8_t s_stage[STAGE_BYTES];
    uint64_t *s_keys = (uint64_t *)s_stage;
    uint32_t *s_vals = (uint32_t *)s_stage;
    uint32_t (*s_wave)[RADIX] = (uint32_t (*)[RADIX])s_stage;
    __shared__ dig_t s_dig[RS_TILE];
    __shared__ uint32_t s_local[RADIX];
    __shared__ uint32_t s_global[RADIX];
    __shared__ uint32_t s_tmp[RS_WAVES];
    __shared__ int s_tile;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = (int)atomicAdd(tile_ctr, 1u);
    for (int i = tid; i < RS_WAVES * RADIX; i += RS_BLOCK) (&s_wave[0][0])[i] = 0;
    __syncthreads();
    const int tile = s_tile;
    int tile_start = tile * RS_TILE, tile_end = n;
    bool first_tile = tile == 0;
    const uint32_t *dbase = digit_base;
    if constexpr (SEG) {
        const int4 ti = seg_tiles[tile];
        tile_start = ti.x;
        tile_end = ti.y;
        first_tile = ti.w != 0;
        dbase = digit_base + (size_t)ti.z * RS_HIST_WORDS;
    }
    const int base = tile_start + wave * (64 * RS_ITEMS);

    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int idx = base + j * 64 + lane;
        const bool ok = idx < tile_end;
        k[j] = ok ? keys_in[idx] : ~0ull;
        v[j] = FIRST ? (uint32_t)idx : (ok ? vals_in[idx] : 0u);
    }
    // Wave-level multisplit: peers = lanes holding the same digit; the rank of
    // an item is the wave's running count of its digit plus the peers below it.
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int idx = base + j * 64 + lane;
        const bool ok = idx < tile_end;
        const uint32_t d = (uint32_t)(k[j] >> shift) & MASK;
        // Per digit bit: bs = all ones when the bit is set, m = the lanes with
        // the bit set; a lane keeps the peers whose bit agrees with its own,
        // peers & ~(m ^ bs), one v_bitop3_b32 per mask half (truth table 0x90
        // for src0 & ~(src1 ^ src2)).
        const uint64_t okm = __ballot(ok);
        uint32_t plo = (uint32_t)okm, phi = (uint32_t)(okm >> 32);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const int32_t bs = __builtin_amdgcn_sbfe((int32_t)d, b, 1);
            const uint64_t m = __ballot(bs < 0);
            plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)m, (uint32_t)bs, 0x90);
            phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(m >> 32), (uint32_t)bs, 0x90);
        }
        const uint64_t peers = ((uint64_t)phi << 32) | plo;
        const uint32_t before = s_wave[wave][d];
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        rank[j] = before + below;
        __builtin_amdgcn_wave_barrier();
        if (ok && below == 0) s_wave[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Per digit (thread = digit, the first RADIX threads): wave offsets, tile
    // count, look-back prefix.
    const uint32_t dg = (uint32_t)tid & MASK;
    const bool dthr = tid < RADIX;
    uint32_t count = 0;
    uint32_t excl = 0;
    if (dthr) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            uint32_t c = s_wave[w][dg];
            s_wave[w][dg] = count;
            count += c;
        }
        const uint64_t tag = (uint64_t)((*epoch * 8u + (uint32_t)pass) & 0x3FFFFFFFu) << 34;
        uint64_t *slot = lookback + (size_t)tile * RADIX + dg;
        if (first_tile) {
            __hip_atomic_store(slot, tag | (2ull << 32) | count, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(slot, tag | (1ull << 32) | count, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            int p = tile - 1;
            while (true) {
                const uint64_t w = __hip_atomic_load(lookback + (size_t)p * RADIX + dg,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((w & ~0x3FFFFFFFFull) != tag || ((w >> 32) & 3u) == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (uint32_t)w;
                if (((w >> 32) & 3u) == 2u) break;
                --p;
            }
            __hip_atomic_store(slot, tag | (2ull << 32) | (uint32_t)(excl + count),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    const uint32_t local = block_excl_scan(count, s_tmp);  // the first RADIX threads hold the digits
    if (dthr) {
        s_local[dg] = local;
        s_global[dg] = dbase[dg] + excl;
    }
    __syncthreads();

    // Tile positions first (the counters share the staging buffer), then the
    // keys into the tile's digit-sorted order in LDS, then the values.
    uint32_t lpos[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int idx = base + j * 64 + lane;
        lpos[j] = 0xFFFFFFFFu;
        if (idx < tile_end) {
            const uint32_t d = (uint32_t)(k[j] >> shift) & MASK;
            lpos[j] = s_local[d] + s_wave[wave][d] + rank[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j)
        if (lpos[j] != 0xFFFFFFFFu) s_keys[lpos[j]] = k[j];
    __syncthreads();
    const int tile_n = min(RS_TILE, tile_end - tile_start);
    for (int i = tid; i < tile_n; i += RS_BLOCK) {
        const uint64_t key = s_keys[i];
        const uint32_t d = (uint32_t)(key >> shift) & MASK;
        s_dig[i] = (dig_t)d;
        keys_out[s_global[d] + ((uint32_t)i - s_local[d])] = key;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j)
        if (lpos[j] != 0xFFFFFFFFu) s_vals[lpos[j]] = v[j];
    __syncthreads();
    for (int i = tid; i < tile_n; i += RS_BLOCK) {
        const uint32_t d = s_dig[i];
        vals_out[s_global[d] + ((uint32_t)i - s_local[d])] = s_vals[i];
    }
}

// ===========================================================================
// Tree assembly: leaf gather (src/hlbvh.hpp:2266-2279), Karras topology
// (src/hlbvh.hpp:773-903) and the bottom-up refit (src/hlbvh.hpp:925-969) in
// one pass over the sorted leaves, plus a second pass for the few nodes that
// span chunks.
//
// The topology is built bottom-up.  split_delta(i) is the common-prefix length
// of leaves i and i+1 over (key, index): clz64 of differing keys, 64 +
// clz32(i ^ (i+1)) for equal keys, -1 outside [0, n-1).  These are the prefix
// lengths the reference's top-down search compares, so the binary radix tree
// -- unique for distinct (key, index) strings -- is the same.  The parent of
// the node over leaves [l, r] is the node split at r (this node is its left
// child) when split_delta(r) > split_delta(l-1), else the node split at l-1
// (right child).  Karras labels follow from that: the root is 0, a left child
// is labelled by its last leaf and a right child by its first leaf, so a
// node's label is known once its parent's side is.  Each node is identified
// by its split until then.  A climber reaching a split posts its far end; the
// second climber to arrive forms the parent from both ends, merges the boxes
// (fmin/fmax for FP, </> for integers: sfc_min/sfc_max, :580-594) and goes on.
// ===========================================================================
__device__ inline int split_delta(const uint64_t *__restrict__ k, int n, int i) {
    if (i < 0 || i + 1 >= n) return -1;
    const uint64_t a = k[i], b = k[i + 1];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)(i + 1)));
}

template <typename T> __device__ inline T smin(T a, T b) {
    if constexpr (CoordTraits<T>::fp) return ffmin(a, b);
    else return (a < b) ? a : b;
}
template <typename T> __device__ inline T smax(T a, T b) {
    if constexpr (CoordTraits<T>::fp) return ffmax(a, b);
    else return (a < b) ? b : a;
}
template <typename T> __device__ inline T ld_agent(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ inline void st_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int ASM_CHUNK = 512;      // leaves per assembly workgroup
static_assert(ASM_CHUNK == SEG_CHUNK && RS_TILE == SEG_TILE && RS_HIST_WORDS == SEG_HIST_WORDS,
              "segmented build tables use the build's tile, chunk and histogram sizes");
constexpr int ASM_CHUNK_MIN = 256;  // sizes the per-chunk top counts in the scratch

// Where the leaves come from: the input particles through the sort
// permutation (build_bvh), caller-sorted arrays (build_tree with leaves), or
// the tree's own leaf slots (build_tree without them).
enum { LEAF_GATHER = 0, LEAF_SORTED = 1, LEAF_INPLACE = 2 };

template <typename T> struct LeafSrc {
    const T *x, *y, *z;
    const uint32_t *id;
    const int8_t *ghost;
    const int32_t *orig;    // LEAF_SORTED only
    const uint32_t *perm;   // LEAF_GATHER only
    const Point4<T> *pts;   // LEAF_GATHER: the points as records (null: x, y, z)
};


// Writes internal node `lab` over leaves [l, r] split at g, with its box
// (agent-scope stores: read by later climbers of the same launch).
template <typename T>
__device__ inline void write_node(const TreeOut<T> &o, int n, int lab, int l, int r, int g,
                                  T b0, T b1, T b2, T b3, T b4, T b5) {
    const int off = n - 1;
    const int li = (l == g) ? off + g : g;
    const int ri = (r == g + 1) ? off + g + 1 : g + 1;
    o.left[lab] = li;
    o.right[lab] = ri;
    o.rfirst[lab] = l;
    o.rlast[lab] = r;
    o.parent[li] = lab;
    o.parent[ri] = lab;
    st_agent(o.mnx + lab, b0); st_agent(o.mxx + lab, b1);
    st_agent(o.mny + lab, b2); st_agent(o.mxy + lab, b3);
    st_agent(o.mnz + lab, b4); st_agent(o.mxz + lab, b5);
}

// Pass 1: workgroup w owns leaves [a, a + C).  Every node whose leaves all lie
// in the chunk is split inside it, and its label lies in it too, so the
// arrival words (one per split), the boxes (by label) and the leaf points sit
// in LDS and the climbs synchronise at workgroup scope; node fields and
// parents are staged in LDS too and leave in coalesced stores.
// A climber whose parent is split at a chunk edge records its node as a chunk
// top and stops.  Afterwards, a split inside the chunk reached only once has
// its other child outside: the child that did arrive is a chunk top too.  The
// chunk tops are disjoint subtrees of the chunk's leaves, so there are at most
// as many as leaves: chunk c lists them in tops[c*C, c*C + count[c]).
// Arrival word of a split: arrivals << 16 | (far end - a + 1) summed.
constexpr int32_t NO_PARENT = INT32_MIN;  // parent not known inside the chunk

// One chunk: leaves [a, a + C) of a tree of n leaves.  Its chunk tops go to
// tops_at[0 ..], their count to *ntops_at.  LEAF_GATHER: perm holds indices
// into the point records that are qoff above the tree's own point indices.
template <typename T, int LEAF, int C>
__device__ __forceinline__ void assemble_chunk(int n, const uint64_t *__restrict__ k,
                                               const LeafSrc<T> &src, const TreeOut<T> &o, int a,
                                               int32_t *__restrict__ tops_at,
                                               uint32_t *__restrict__ ntops_at, uint32_t qoff) {
    static_assert(C <= 1024, "arrival words hold chunk offsets in 16 bits");
    __shared__ T spt[3][C];
    __shared__ T sbox[6][C];
    __shared__ int32_t sl[C], sr[C], sfirst[C], slast[C];  // by label - a; sfirst < 0: not here
    __shared__ int32_t spar_in[C], spar_leaf[C];           // by label - a, by leaf - a
    __shared__ uint32_t sarr[C];
    __shared__ int8_t sdel[C + 1];  // sdel[j] = split_delta(a - 1 + j), in [-1, 96]
    __shared__ uint32_t s_ntop;
    const int t = threadIdx.x;
    const int i = a + t;
    const int off = n - 1;
    T px = T(0), py = T(0), pz = T(0);
    if (i < n) {
        const int leaf = off + i;
        if constexpr (LEAF == LEAF_GATHER) {
            const uint32_t qg = src.perm[i];
            const uint32_t q = qg - qoff;
            if (src.pts) {
                const Point4<T> p4 = src.pts[qg];
                px = p4.v[0]; py = p4.v[1]; pz = p4.v[2];
            } else {
                px = src.x[q]; py = src.y[q]; pz = src.z[q];
            }
            o.id[leaf] = src.id ? src.id[q] : q;
            o.ghost[leaf] = src.ghost ? src.ghost[q] : (int8_t)0;
            o.orig[leaf] = (int32_t)q;
        } else if constexpr (LEAF == LEAF_SORTED) {
            px = src.x[i]; py = src.y[i]; pz = src.z[i];
            if (src.id) o.id[leaf] = src.id[i];
            if (src.ghost) o.ghost[leaf] = src.ghost[i];
            if (src.orig) o.orig[leaf] = src.orig[i];
        } else {
            px = o.mnx[leaf]; py = o.mny[leaf]; pz = o.mnz[leaf];
        }
        if constexpr (LEAF != LEAF_INPLACE) {
            o.mnx[leaf] = px; o.mxx[leaf] = px;
            o.mny[leaf] = py; o.mxy[leaf] = py;
            o.mnz[leaf] = pz; o.mxz[leaf] = pz;
        }
        spt[0][t] = px; spt[1][t] = py; spt[2][t] = pz;
    }
    sdel[t + 1] = (int8_t)split_delta(k, n, i);
    if (t == 0) {
        sdel[0] = (int8_t)split_delta(k, n, a - 1);
        s_ntop = 0u;
    }
    sarr[t] = 0u;
    sfirst[t] = -1;
    spar_in[t] = NO_PARENT;
    spar_leaf[t] = NO_PARENT;
    __syncthreads();

    if (i < n) {
        int l = i, r = i, g = -1;
        T b0 = px, b1 = px, b2 = py, b3 = py, b4 = pz, b5 = pz;
        while (true) {
            const bool root = (l == 0 && r == off);
            const bool up_right = !root && sdel[r - a + 1] > sdel[l - a];
            const int lab = (l == r) ? off + l : (root ? 0 : (up_right ? r : l));
            if (l < r) {  // this climber completed internal node `lab`, split at g
                const int j = lab - a;
                sl[j] = (l == g) ? off + g : g;
                sr[j] = (r == g + 1) ? off + g + 1 : g + 1;
                sfirst[j] = l;
                slast[j] = r;
                if (l == g) spar_leaf[g - a] = lab;
                else spar_in[g - a] = lab;
                if (r == g + 1) spar_leaf[g + 1 - a] = lab;
                else spar_in[g + 1 - a] = lab;
                sbox[0][j] = b0; sbox[1][j] = b1; sbox[2][j] = b2;
                sbox[3][j] = b3; sbox[4][j] = b4; sbox[5][j] = b5;
            }
            if (root) {
                if (l < r) spar_in[0] = -1;
                else spar_leaf[0] = -1;  // a single leaf is the root
                break;
            }
            const int p = up_right ? r : l - 1;
            if (p < a || p > a + C - 2) {
                tops_at[(int)atomicAdd(&s_ntop, 1u)] = lab;
                break;
            }
            const uint32_t far = (uint32_t)((up_right ? l : r) - a + 1);
            const uint32_t prev = __hip_atomic_fetch_add(&sarr[p - a], (1u << 16) | far,
                                                         __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prev == 0u) break;
            const int sf = (int)(prev & 0xFFFFu) - 1 + a;
            const bool sleaf = up_right ? (sf == p + 1) : (sf == p);
            const int sj = (up_right ? p + 1 : p) - a;
            if (sleaf) {
                const T x = spt[0][sj], y = spt[1][sj], z = spt[2][sj];
                b0 = smin(b0, x); b1 = smax(b1, x);
                b2 = smin(b2, y); b3 = smax(b3, y);
                b4 = smin(b4, z); b5 = smax(b5, z);
            } else {
                b0 = smin(b0, sbox[0][sj]); b1 = smax(b1, sbox[1][sj]);
                b2 = smin(b2, sbox[2][sj]); b3 = smax(b3, sbox[3][sj]);
                b4 = smin(b4, sbox[4][sj]); b5 = smax(b5, sbox[5][sj]);
            }
            if (up_right) r = sf;
            else l = sf;
            g = p;
        }
    }
    __syncthreads();
    // Chunk tops left by splits reached once.
    {
        const int p = a + t;
        if (t < C - 1 && p < off) {
            const uint32_t v = sarr[t];
            if ((v >> 16) == 1u) {
                const int f = (int)(v & 0xFFFFu) - 1 + a;
                const int lab = (f <= p) ? ((f == p) ? off + p : p)
                                         : ((f == p + 1) ? off + p + 1 : p + 1);
                tops_at[(int)atomicAdd(&s_ntop, 1u)] = lab;
            }
        }
    }
    // Coalesced write-back of what was completed here.
    if (i < off && sfirst[t] >= 0) {
        o.left[i] = sl[t];
        o.right[i] = sr[t];
        o.rfirst[i] = sfirst[t];
        o.rlast[i] = slast[t];
        o.mnx[i] = sbox[0][t]; o.mxx[i] = sbox[1][t];
        o.mny[i] = sbox[2][t]; o.mxy[i] = sbox[3][t];
        o.mnz[i] = sbox[4][t]; o.mxz[i] = sbox[5][t];
    }
    if (i < off && spar_in[t] != NO_PARENT) o.parent[i] = spar_in[t];
    if (i < n && spar_leaf[t] != NO_PARENT) o.parent[off + i] = spar_leaf[t];
    __syncthreads();
    if (t == 0 && a < n) *ntops_at = s_ntop;
}

// Pass 1 of one tree: chunk c = leaves [c C, c C + C), XCD-ordered so that the
// gather's reads of one cluster go through one L2.
template <typename T, int LEAF, int C>
__global__ __launch_bounds__(C) void assemble_kernel(int n, const uint64_t *__restrict__ k,
                                                     LeafSrc<T> src, TreeOut<T> o,
                                                     int32_t *__restrict__ tops,
                                                     uint32_t *__restrict__ chunk_tops) {
    const int chunk = xcd_block();
    assemble_chunk<T, LEAF, C>(n, k, src, o, chunk * C, tops + (size_t)chunk * C,
                               chunk_tops + chunk, 0u);
}

// Pass 2: the chunk tops climb through the nodes that span chunks.  Arrival
// words are global, one per split, tagged with a device-side build counter so
// that no clearing pass is needed: word = epoch << 32 | far end, swapped in
// atomically; a swapped-out word of an older epoch means "first".  Boxes are
// written with agent-scope stores (write-through to the device coherence
// point), each climber waits for its own stores (s_waitcnt vmcnt(0)) before
// its swap, and the second arrival reads the sibling with agent-scope loads
// after the swap returned -- the agent-scope release/acquire pair without the
// L2 write-back / invalidate that only non-atomic accesses need.  Pass 1's
// writes are visible after the kernel boundary.
// The tops of one chunk (tops_at[0 .. count)) climb, one lane each.
template <typename T>
__device__ __forceinline__ void climb_tops(int n, const uint64_t *__restrict__ k,
                                           const TreeOut<T> &o, const int32_t *__restrict__ tops_at,
                                           int count, uint64_t *__restrict__ words, uint64_t ep,
                                           int lane) {
    const int off = n - 1;
    for (int q = lane; q < count; q += 64) {
        const int lab0 = tops_at[q];
        if (lab0 < 0 || lab0 > 2 * off) continue;  // never true for a list this build wrote
        int l, r;
        if (lab0 >= off) {
            l = r = lab0 - off;
        } else {
            l = o.rfirst[lab0];
            r = o.rlast[lab0];
        }
        if (l < 0 || r > off || l > r) continue;
        T b0 = o.mnx[lab0], b1 = o.mxx[lab0], b2 = o.mny[lab0];
        T b3 = o.mxy[lab0], b4 = o.mnz[lab0], b5 = o.mxz[lab0];
        int dl = split_delta(k, n, l - 1), dr = split_delta(k, n, r);
        while (!(l == 0 && r == off)) {
            const bool up_right = dr > dl;
            const int p = up_right ? r : l - 1;
            const uint64_t mine = ep | (uint32_t)(up_right ? l : r);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            const uint64_t prev =
                __hip_atomic_exchange(&words[p], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((prev & 0xFFFFFFFF00000000ull) != ep) break;
            asm volatile("" ::: "memory");
            const int sf = (int)(uint32_t)prev;
            const int sib = up_right ? ((sf == p + 1) ? off + p + 1 : p + 1)
                                     : ((sf == p) ? off + p : p);
            b0 = smin(b0, ld_agent(o.mnx + sib)); b1 = smax(b1, ld_agent(o.mxx + sib));
            b2 = smin(b2, ld_agent(o.mny + sib)); b3 = smax(b3, ld_agent(o.mxy + sib));
            b4 = smin(b4, ld_agent(o.mnz + sib)); b5 = smax(b5, ld_agent(o.mxz + sib));
            if (up_right) r = sf;
            else l = sf;
            const bool root = (l == 0 && r == off);
            int lab = 0;
            if (!root) {
                dl = split_delta(k, n, l - 1);
                dr = split_delta(k, n, r);
                lab = (dr > dl) ? r : l;
            }
            write_node<T>(o, n, lab, l, r, p, b0, b1, b2, b3, b4, b5);
            if (root) o.parent[0] = -1;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void assemble_top_kernel(int n, const uint64_t *__restrict__ k,
                                                           TreeOut<T> o,
                                                           const int32_t *__restrict__ tops,
                                                           const uint32_t *__restrict__ chunk_tops,
                                                           int chunks, int chunk_size,
                                                           uint64_t *__restrict__ words,
                                                           const uint32_t *__restrict__ epoch) {
    const uint64_t ep = (uint64_t)*epoch << 32;
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    // one wave per chunk, one lane per chunk top; XCD x takes a contiguous run
    // of chunks, so the first climbs between neighbouring chunks stay on it
    for (int c = xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6); c < chunks; c += waves)
        climb_tops<T>(n, k, o, tops + (size_t)c * chunk_size, (int)chunk_tops[c], words, ep, lane);
}

// Starts a topology pass: advances the arrival words' epoch.
__global__ void topo_begin_kernel(uint32_t *epoch) {
    if (threadIdx.x == 0) *epoch = *epoch + 1u;
}

// ===========================================================================
// Host orchestration
// ===========================================================================
static inline unsigned grid_for(size_t n, int block, unsigned cap = 0) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (cap && g > cap) g = cap;
    return (unsigned)g;
}

static int scratch_ensure(Scratch &S, size_t n) {
    if (n <= S.cap_n && S.block) return FT_OK;
    if (S.block) {
        FT_HIP(hipDeviceSynchronize());  // earlier launches may still use the old block
        FT_HIP(hipFree(S.block));
        S.block = nullptr;
    }
    const size_t cap = n + n / 2 + 1024;
    const size_t tiles = (cap + RS_TILE_MIN - 1) / RS_TILE_MIN + 1;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_keys_in = off; off += al(cap * 8);
    const size_t o_keys_a = off; off += al(cap * 8);
    const size_t o_keys_b = off; off += al(cap * 8);
    const size_t o_vals_a = off; off += al(cap * 4);
    const size_t o_vals_b = off; off += al(cap * 4);
    const size_t o_perm = off; off += al(cap * 4);
    const size_t o_words = off; off += al(cap * 8);
    const size_t o_tops = off; off += al(cap * 4);
    const size_t o_hist = off; off += al(RS_HIST_WORDS * 4 * 2);
    const size_t o_look = off; off += al(tiles * RS_MAX_RADIX * 8);
    const size_t o_tile = off; off += al(64);
    const size_t o_epoch = off; off += al(64);
    const size_t o_ctops = off; off += al(4 * (cap / ASM_CHUNK_MIN + 2));
    const size_t o_tepoch = off; off += al(64);
    const size_t o_bbox = off; off += al(16 * 8);
    const size_t o_part = off; off += al(1024 * 6 * 8);
    const size_t o_pts = off; off += al(cap * 32);
    void *blk = nullptr;
    if (hipMalloc(&blk, off) != hipSuccess) return fail(FT_ENOMEM, "scratch allocation failed");
    // One-time clear: look-back and arrival words carry an epoch, so they never
    // need clearing again.  hipMemset runs on the null stream, which callers'
    // non-blocking streams do not wait for: finish it before any of them can
    // use the block.
    FT_HIP(hipMemset(blk, 0, off));
    FT_HIP(hipDeviceSynchronize());
    char *b = (char *)blk;
    S.block = blk;
    S.block_bytes = off;
    S.cap_n = cap;
    S.keys_in = (uint64_t *)(b + o_keys_in);
    S.keys_a = (uint64_t *)(b + o_keys_a);
    S.keys_b = (uint64_t *)(b + o_keys_b);
    S.vals_a = (uint32_t *)(b + o_vals_a);
    S.vals_b = (uint32_t *)(b + o_vals_b);
    S.perm = (uint32_t *)(b + o_perm);
    S.split_words = (uint64_t *)(b + o_words);
    S.tops = (int32_t *)(b + o_tops);
    S.chunk_tops = (uint32_t *)(b + o_ctops);
    S.topo_epoch = (uint32_t *)(b + o_tepoch);
    S.hist = (uint32_t *)(b + o_hist);
    S.lookback = (uint64_t *)(b + o_look);
    S.lookback_words = tiles * RS_MAX_RADIX;
    S.tile_ctr = (uint32_t *)(b + o_tile);
    S.epoch = (uint32_t *)(b + o_epoch);
    S.bbox_dev = (void *)(b + o_bbox);
    S.partials = (void *)(b + o_part);
    S.points = (void *)(b + o_pts);
    return FT_OK;
}

int scratch_ensure_public(Scratch &S, size_t n) { return scratch_ensure(S, n); }

static void prof_mark(Scratch &S, hipStream_t s, int idx) {
    if (!S.profiling) return;
    if (!S.ev[idx]) hipEventCreate(&S.ev[idx]);
    hipEventRecord(S.ev[idx], s);
}

template <typename T>
static int launch_bbox(hipStream_t s, Scratch &S, const T *x, const T *y, const T *z, size_t n,
                       const T *given6) {
    BboxDev<T> *bb = (BboxDev<T> *)S.bbox_dev;
    if (given6) {
        hipLaunchKernelGGL(bbox_final_kernel<T>, dim3(1), dim3(64), 0, s, (const T *)nullptr, 0, 1,
                           given6[0], given6[1], given6[2], given6[3], given6[4], given6[5], bb);
    } else {
        const unsigned g = grid_for(n, 256, 1024);
        hipLaunchKernelGGL(bbox_partial_kernel<T>, dim3(g), dim3(256), 0, s, x, y, z, n,
                           (T *)S.partials);
        hipLaunchKernelGGL(bbox_final_kernel<T>, dim3(1), dim3(64), 0, s, (const T *)S.partials,
                           (int)g, 0, T(0), T(0), T(0), T(0), T(0), T(0), bb);
    }
    FT_CHECK_LAUNCH();
    return FT_OK;
}

template <typename T>
static int launch_encode(hipStream_t s, Scratch &S, const ft_config &cfg, const T *x, const T *y,
                         const T *z, size_t n, uint64_t *keys, uint32_t *hist = nullptr,
                         int passes = 0, int rbits = 8, Point4<T> *pts = nullptr) {
    const unsigned g = grid_for(n, 256, 4096);
    hipLaunchKernelGGL(encode_kernel<T>, dim3(g), dim3(256), 0, s, x, y, z, n, cfg.bits_per_dim,
                       cfg.sfc, (const BboxDev<T> *)S.bbox_dev, keys, passes, rbits, hist, pts);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// Sorts keys_in (n keys of `key_bits` significant bits) into sorted_out, and
// the permutation into S.perm, in passes of S.radix_bits-bit digits.
static inline int radix_passes(int key_bits, int rbits) {
    const int p = (key_bits + rbits - 1) / rbits;
    return p < 1 ? 1 : p;
}
// The digit width a sort of key_bits-bit keys uses: the scratch's setting, or
// 9 bits when that saves a pass (63-bit keys: 7 passes instead of 8).
static inline int radix_width(const Scratch &S, int key_bits) {
    if (S.radix_bits) return S.radix_bits;
    return radix_passes(key_bits, 9) < radix_passes(key_bits, 8) ? 9 : 8;
}

template <int BITS>
static void launch_sweeps(hipStream_t s, Scratch &S, const uint64_t *keys_in, size_t n,
                          int passes, const uint32_t *base, uint64_t *sorted_out) {
    const unsigned tiles = (unsigned)((n + RS_TILE - 1) / RS_TILE);
    const uint64_t *kin = keys_in;
    const uint32_t *vin = nullptr;
    for (int p = 0; p < passes; ++p) {
        const bool last = (p == passes - 1);
        uint64_t *kout = last ? sorted_out : ((p & 1) ? S.keys_b : S.keys_a);
        uint32_t *vout = last ? S.perm : ((p & 1) ? S.vals_b : S.vals_a);
        if (p == 0)
            hipLaunchKernelGGL((onesweep_kernel<true, BITS>), dim3(tiles), dim3(RS_BLOCK), 0, s,
                               kin, vin, kout, vout, (int)n, BITS * p, base + (p << BITS),
                               S.lookback, (const uint32_t *)S.epoch, p, S.tile_ctr + p);
        else
            hipLaunchKernelGGL((onesweep_kernel<false, BITS>), dim3(tiles), dim3(RS_BLOCK), 0, s,
                               kin, vin, kout, vout, (int)n, BITS * p, base + (p << BITS),
                               S.lookback, (const uint32_t *)S.epoch, p, S.tile_ctr + p);
        kin = kout;
        vin = vout;
    }
}

static int radix_sort(hipStream_t s, Scratch &S, const uint64_t *keys_in, size_t n, int key_bits,
                      uint64_t *sorted_out, bool hist_ready) {
    const int rbits = radix_width(S, key_bits);
    const int passes = radix_passes(key_bits, rbits);
    if (passes > RS_MAX_PASSES) return fail(FT_EINVAL, "radix sort: too many passes");
    uint32_t *base = S.hist + RS_HIST_WORDS;
    if (!hist_ready) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(RS_HIST_WORDS, 256)), dim3(256), 0, s,
                           S.hist, 0u, (size_t)RS_HIST_WORDS);
        hipLaunchKernelGGL(radix_hist_kernel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s,
                           keys_in, n, passes, rbits, S.hist);
    }
    hipLaunchKernelGGL(radix_base_kernel, dim3(1), dim3(1 << rbits), 0, s, (const uint32_t *)S.hist,
                       passes, rbits, base, S.tile_ctr, S.epoch);
    FT_CHECK_LAUNCH();
    if (rbits == 9) launch_sweeps<9>(s, S, keys_in, n, passes, base, sorted_out);
    else launch_sweeps<8>(s, S, keys_in, n, passes, base, sorted_out);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// Stable sort of n keys of key_bits significant bits for other modules:
// sorted keys to sorted_out, the permutation (sorted position -> input index)
// to S.perm.
int radix_sort_u64(hipStream_t s, Scratch &S, const uint64_t *keys_in, size_t n,
                   uint64_t *sorted_out, int key_bits) {
    FT_TRY(scratch_ensure(S, n));
    return radix_sort(s, S, keys_in, n, key_bits, sorted_out, false);
}

// Assembles the tree over t->sorted_keys: leaves from `src` (LEAF mode),
// topology, ranges and boxes.  Pass 2 runs only when the tree spans chunks.
// 512-leaf chunks: three workgroups per CU with 8-byte coordinates.
template <typename T, int LEAF>
static int launch_assemble(hipStream_t s, ft_tree *t, Scratch &S, const LeafSrc<T> &src) {
    const int n = (int)t->num_leaves;
    if (n < 1) return FT_OK;
    const TreeOut<T> o{(T *)t->min_x,      (T *)t->max_x,       (T *)t->min_y, (T *)t->max_y,
                       (T *)t->min_z,      (T *)t->max_z,       t->left_child, t->right_child,
                       t->parent,          t->range_first,      t->range_last, t->id,
                       t->is_ghost,        t->orig_idx};
    const uint64_t *keys = (const uint64_t *)t->sorted_keys;
    constexpr int C = ASM_CHUNK;
    hipLaunchKernelGGL(topo_begin_kernel, dim3(1), dim3(64), 0, s, S.topo_epoch);
    const unsigned chunks = (unsigned)((n + C - 1) / C);
    hipLaunchKernelGGL((assemble_kernel<T, LEAF, C>), dim3(xcd_grid(chunks)), dim3(C), 0, s,
                       n, keys, src, o, S.tops, S.chunk_tops);
    FT_CHECK_LAUNCH();
    prof_mark(S, s, 4);
    if (chunks > 1) {
        hipLaunchKernelGGL(assemble_top_kernel<T>, dim3(xcd_grid((chunks + 3) / 4)), dim3(256), 0, s, n, keys,
                           o, (const int32_t *)S.tops, (const uint32_t *)S.chunk_tops, (int)chunks,
                           C, S.split_words, (const uint32_t *)S.topo_epoch);
        FT_CHECK_LAUNCH();
    }
    return FT_OK;
}

template <typename T>
int build_bvh_t(hipStream_t s, ft_tree *t, Scratch &S, const T *x, const T *y, const T *z,
                const uint32_t *id, const int8_t *ghost, size_t n, const T *bbox6,
                const uint64_t *keys) {
    FT_TRY(scratch_ensure(S, n));
    const ft_config &cfg = t->config;
    prof_mark(S, s, 0);
    const uint64_t *kin = keys;
    const int key_bits = 3 * cfg.bits_per_dim;
    if (!keys) {
        if constexpr (CoordTraits<T>::fp) FT_TRY(launch_bbox<T>(s, S, x, y, z, n, bbox6));
        prof_mark(S, s, 1);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(RS_HIST_WORDS, 256)), dim3(256), 0, s,
                           S.hist, 0u, (size_t)RS_HIST_WORDS);
        const int rbits = radix_width(S, key_bits);
        FT_TRY(launch_encode<T>(s, S, cfg, x, y, z, n, S.keys_in, S.hist,
                                radix_passes(key_bits, rbits), rbits, (Point4<T> *)S.points));
        kin = S.keys_in;
    } else {
        prof_mark(S, s, 1);
    }
    prof_mark(S, s, 2);
    FT_TRY(radix_sort(s, S, kin, n, key_bits, t->sorted_keys, keys == nullptr));
    prof_mark(S, s, 3);
    // keys encoded here: the encode pass also wrote the points as records
    LeafSrc<T> src{x, y, z, id, ghost, nullptr, (const uint32_t *)S.perm,
                   keys ? nullptr : (const Point4<T> *)S.points};
    FT_TRY((launch_assemble<T, LEAF_GATHER>(s, t, S, src)));
    prof_mark(S, s, 5);
    if (S.profiling) S.n_stages = 5;
    return FT_OK;
}

int collect_profile(Scratch &S) {
    if (!S.profiling || S.n_stages == 0) return FT_OK;
    FT_HIP(hipEventSynchronize(S.ev[S.n_stages]));
    for (int i = 0; i < S.n_stages; ++i) {
        float ms = 0.f;
        FT_HIP(hipEventElapsedTime(&ms, S.ev[i], S.ev[i + 1]));
        S.stage_ms[i] = ms;
    }
    return FT_OK;
}

template int build_bvh_t<float>(hipStream_t, ft_tree *, Scratch &, const float *, const float *,
                                const float *, const uint32_t *, const int8_t *, size_t,
                                const float *, const uint64_t *);
template int build_bvh_t<double>(hipStream_t, ft_tree *, Scratch &, const double *,
                                 const double *, const double *, const uint32_t *,
                                 const int8_t *, size_t, const double *, const uint64_t *);
template int build_bvh_t<uint32_t>(hipStream_t, ft_tree *, Scratch &, const uint32_t *,
                                   const uint32_t *, const uint32_t *, const uint32_t *,
                                   const int8_t *, size_t, const uint32_t *, const uint64_t *);
template int build_bvh_t<uint64_t>(hipStream_t, ft_tree *, Scratch &, const uint64_t *,
                                   const uint64_t *, const uint64_t *, const uint32_t *,
                                   const int8_t *, size_t, const uint64_t *, const uint64_t *);

// compute_bbox: device reduction, result copied to the host (blocking, as the
// reference reads the shared-USM result on the host).
template <typename T>
int compute_bbox_t(hipStream_t s, Scratch &S, const T *x, const T *y, const T *z, size_t n,
                   T *out6) {
    FT_TRY(scratch_ensure(S, 1));
    FT_TRY(launch_bbox<T>(s, S, x, y, z, n, nullptr));
    BboxDev<T> hb;
    FT_HIP(hipMemcpyAsync(&hb, S.bbox_dev, sizeof(hb), hipMemcpyDeviceToHost, s));
    FT_HIP(hipStreamSynchronize(s));
    for (int a = 0; a < 3; ++a) {
        out6[2 * a] = hb.lo[a];
        out6[2 * a + 1] = hb.hi[a];
    }
    return FT_OK;
}
template int compute_bbox_t<float>(hipStream_t, Scratch &, const float *, const float *,
                                   const float *, size_t, float *);
template int compute_bbox_t<double>(hipStream_t, Scratch &, const double *, const double *,
                                    const double *, size_t, double *);
template int compute_bbox_t<uint32_t>(hipStream_t, Scratch &, const uint32_t *,
                                      const uint32_t *, const uint32_t *, size_t, uint32_t *);
template int compute_bbox_t<uint64_t>(hipStream_t, Scratch &, const uint64_t *,
                                      const uint64_t *, const uint64_t *, size_t, uint64_t *);

// sfc_encode with a caller box (FP) into device keys.
template <typename T>
int sfc_encode_t(hipStream_t s, Scratch &S, const ft_config &cfg, const T *x, const T *y,
                 const T *z, size_t n, uint64_t *keys, const T *bbox6) {
    FT_TRY(scratch_ensure(S, 1));
    if constexpr (CoordTraits<T>::fp) {
        if (!bbox6) return fail(FT_EINVAL, "sfc_encode: FP coordinates need a bounding box");
        FT_TRY(launch_bbox<T>(s, S, x, y, z, n, bbox6));
    }
    return launch_encode<T>(s, S, cfg, x, y, z, n, keys);
}
template int sfc_encode_t<float>(hipStream_t, Scratch &, const ft_config &, const float *,
                                 const float *, const float *, size_t, uint64_t *, const float *);
template int sfc_encode_t<double>(hipStream_t, Scratch &, const ft_config &, const double *,
                                  const double *, const double *, size_t, uint64_t *,
                                  const double *);
template int sfc_encode_t<uint32_t>(hipStream_t, Scratch &, const ft_config &, const uint32_t *,
                                    const uint32_t *, const uint32_t *, size_t, uint64_t *,
                                    const uint32_t *);
template int sfc_encode_t<uint64_t>(hipStream_t, Scratch &, const ft_config &, const uint64_t *,
                                    const uint64_t *, const uint64_t *, size_t, uint64_t *,
                                    const uint64_t *);

// build_tree over caller-sorted keys (+ optional sorted leaves).
template <typename T>
int build_tree_t(hipStream_t s, ft_tree *t, Scratch &S, const uint64_t *sorted_keys, const T *sx,
                 const T *sy, const T *sz, const uint32_t *sid, const int8_t *sghost,
                 const int32_t *sorig) {
    const size_t n = t->num_leaves;
    FT_TRY(scratch_ensure(S, n));
    if (sorted_keys != t->sorted_keys)
        FT_HIP(hipMemcpyAsync(t->sorted_keys, sorted_keys, n * 8, hipMemcpyDeviceToDevice, s));
    if (sx && sy && sz) {
        LeafSrc<T> src{sx, sy, sz, sid, sghost, sorig, nullptr};
        return launch_assemble<T, LEAF_SORTED>(s, t, S, src);
    }
    LeafSrc<T> src{};
    return launch_assemble<T, LEAF_INPLACE>(s, t, S, src);
}
template int build_tree_t<float>(hipStream_t, ft_tree *, Scratch &, const uint64_t *,
                                 const float *, const float *, const float *, const uint32_t *,
                                 const int8_t *, const int32_t *);
template int build_tree_t<double>(hipStream_t, ft_tree *, Scratch &, const uint64_t *,
                                  const double *, const double *, const double *,
                                  const uint32_t *, const int8_t *, const int32_t *);
template int build_tree_t<uint32_t>(hipStream_t, ft_tree *, Scratch &, const uint64_t *,
                                    const uint32_t *, const uint32_t *, const uint32_t *,
                                    const uint32_t *, const int8_t *, const int32_t *);
template int build_tree_t<uint64_t>(hipStream_t, ft_tree *, Scratch &, const uint64_t *,
                                    const uint64_t *, const uint64_t *, const uint64_t *,
                                    const uint32_t *, const int8_t *, const int32_t *);

// ===========================================================================
// Segmented build (the batched step): every tree of a batch in one launch of
// each kernel.  Tree t's points are indices [off, off + n) of the combined
// numbering; its sort tiles and assembly chunks are its own (tables built on
// the host), its digit histograms and bases are its own (the bases include
// off, so the sort writes each tree's keys into its own range), and the
// look-back of a tile stops at its tree's first tile.  Keys, permutation,
// topology and boxes are those of build_bvh on the tree alone: the sort is
// the same stable sort per tree, and the assembly the same per chunk.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void seg_bbox_kernel(const SegTree<T> *__restrict__ tr,
                                                       const int4 *__restrict__ tiles,
                                                       T *__restrict__ partials) {
    const int4 ti = tiles[blockIdx.x];
    const SegTree<T> st = tr[ti.z];
    const int i0 = ti.x - st.off, i1 = ti.y - st.off;
    T v[6];
    {
        const T f = st.x[i0], g = st.y[i0], h = st.z[i0];
        v[0] = f; v[1] = f; v[2] = g; v[3] = g; v[4] = h; v[5] = h;
    }
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
        const T a = st.x[i], b = st.y[i], c = st.z[i];
        v[0] = tmin(v[0], a); v[1] = tmax(v[1], a);
        v[2] = tmin(v[2], b); v[3] = tmax(v[3], b);
        v[4] = tmin(v[4], c); v[5] = tmax(v[5], c);
    }
    __shared__ T red[4][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const T r = (a & 1) ? wave_max(v[a]) : wave_min(v[a]);
        if (lane == 0) red[w][a] = r;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        T r = red[0][a];
        for (int q = 1; q < 4; ++q) r = (a & 1) ? tmax(r, red[q][a]) : tmin(r, red[q][a]);
        partials[(size_t)blockIdx.x * 6 + a] = r;
    }
}

// One block per tree: its tiles' partials -> its box record (as bbox_final_kernel).
template <typename T>
__global__ __launch_bounds__(64) void seg_bbox_final_kernel(const SegTree<T> *__restrict__ tr,
                                                            const T *__restrict__ partials,
                                                            BboxDev<T> *__restrict__ out) {
    const SegTree<T> st = tr[blockIdx.x];
    if (threadIdx.x != 0 || st.n <= 0) return;
    T v[6];
    for (int a = 0; a < 6; ++a) v[a] = partials[(size_t)st.tile0 * 6 + a];
    for (int b = 1; b < st.ntiles; ++b)
        for (int a = 0; a < 6; ++a) {
            const T pv = partials[(size_t)(st.tile0 + b) * 6 + a];
            v[a] = (a & 1) ? tmax(v[a], pv) : tmin(v[a], pv);
        }
    BboxDev<T> &o = out[blockIdx.x];
    for (int a = 0; a < 3; ++a) {
        o.lo[a] = v[2 * a];
        o.hi[a] = v[2 * a + 1];
        const T d = v[2 * a + 1] - v[2 * a];
        o.inv[a] = (d == T(0)) ? T(0) : (T(1) / d);
    }
}

// One block per tile: keys, point records and the tile's tree's digit histograms.
template <typename T>
__global__ __launch_bounds__(256) void seg_encode_kernel(const SegTree<T> *__restrict__ tr,
                                                         const int4 *__restrict__ tiles, int B,
                                                         int sfc, const BboxDev<T> *__restrict__ bb,
                                                         uint64_t *__restrict__ keys, int passes,
                                                         int rbits, uint32_t *__restrict__ hist,
                                                         Point4<T> *__restrict__ pts) {
    __shared__ uint16_t tab[384];
    __shared__ uint32_t h[ENC_HIST_WORDS];
    for (int i = threadIdx.x; i < 384; i += blockDim.x) tab[i] = c_ph_table[i];
    for (int i = threadIdx.x; i < ENC_HIST_WORDS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int4 ti = tiles[blockIdx.x];
    const SegTree<T> st = tr[ti.z];
    T lo[3] = {T(0), T(0), T(0)}, inv[3] = {T(0), T(0), T(0)};
    if constexpr (CoordTraits<T>::fp) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = bb[ti.z].lo[a];
            inv[a] = bb[ti.z].inv[a];
        }
    }
    for (int g = ti.x + (int)threadIdx.x; g < ti.y; g += blockDim.x) {
        const int i = g - st.off;
        const T p[3] = {st.x[i], st.y[i], st.z[i]};
        pts[g] = Point4<T>{{p[0], p[1], p[2], T(0)}};
        const uint64_t key = encode_point<T>(p, B, sfc, lo, inv, tab);
        keys[g] = key;
        hist_add_key(h, key, passes, rbits);
    }
    __syncthreads();
    uint32_t *ht = hist + (size_t)ti.z * RS_HIST_WORDS;
    for (int i = threadIdx.x; i < (passes << rbits); i += blockDim.x) {
        const uint32_t v = h[i];
        if (v) atomicAdd(&ht[i], v);
    }
}

// One block per tree: digit bases = the tree's first index + the exclusive
// scan of its histogram, per pass.  Block 0 also resets the tile counters
// and advances the sort epoch (as radix_base_kernel).
template <typename T>
__global__ __launch_bounds__(RS_MAX_RADIX) void seg_radix_base_kernel(
    const SegTree<T> *__restrict__ tr, const uint32_t *__restrict__ hist, int passes, int rbits,
    uint32_t *__restrict__ base, uint32_t *__restrict__ tile_ctr, uint32_t *__restrict__ epoch) {
    __shared__ uint32_t tmp[RS_MAX_RADIX / 64];
    const int t = blockIdx.x;
    const uint32_t off = (uint32_t)tr[t].off;
    for (int p = 0; p < passes; ++p) {
        const size_t w = (size_t)t * RS_HIST_WORDS + ((size_t)p << rbits) + threadIdx.x;
        base[w] = off + block_excl_scan(hist[w], tmp);
    }
    if (t == 0) {
        if (threadIdx.x < RS_MAX_PASSES) tile_ctr[threadIdx.x] = 0;
        if (threadIdx.x == 0) {
            const uint32_t e = *epoch + 1u;
            *epoch = (e >= (1u << 27)) ? 1u : e;
        }
    }
}

// Pass 1 of every tree: chunk (tree, first leaf) from the chunk table.  Each
// chunk also copies its sorted keys into its tree's sorted_keys.
template <typename T, int C>
__global__ __launch_bounds__(C) void seg_assemble_kernel(const SegTree<T> *__restrict__ tr,
                                                         const int2 *__restrict__ chunks,
                                                         int nchunks,
                                                         const uint64_t *__restrict__ sorted,
                                                         const uint32_t *__restrict__ perm,
                                                         const Point4<T> *__restrict__ pts,
                                                         int32_t *__restrict__ tops,
                                                         uint32_t *__restrict__ chunk_tops) {
    const int chunk = xcd_block();
    if (chunk >= nchunks) return;  // padding block of the XCD-ordered grid
    const int2 ci = chunks[chunk];
    const SegTree<T> st = tr[ci.x];
    const uint64_t *k = sorted + st.off;
    const int a = ci.y;
    if (a + (int)threadIdx.x < st.n) st.sorted_keys[a + threadIdx.x] = k[a + threadIdx.x];
    LeafSrc<T> src{st.x, st.y, st.z, nullptr, nullptr, nullptr, perm + st.off, pts};
    assemble_chunk<T, LEAF_GATHER, C>(st.n, k, src, st.o, a, tops + (size_t)chunk * C,
                                      chunk_tops + chunk, (uint32_t)st.off);
}

// Pass 2 of every tree: each wave takes chunks and climbs their tops within
// their trees; the arrival words of tree t are words[off ..].
template <typename T>
__global__ __launch_bounds__(256) void seg_top_kernel(const SegTree<T> *__restrict__ tr,
                                                      const int2 *__restrict__ chunks,
                                                      int nchunks,
                                                      const uint64_t *__restrict__ sorted,
                                                      const int32_t *__restrict__ tops,
                                                      const uint32_t *__restrict__ chunk_tops,
                                                      uint64_t *__restrict__ words,
                                                      const uint32_t *__restrict__ epoch) {
    const uint64_t ep = (uint64_t)*epoch << 32;
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int c = xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6); c < nchunks; c += waves) {
        const int cnt = (int)chunk_tops[c];
        if (cnt == 0) continue;
        const SegTree<T> st = tr[chunks[c].x];
        climb_tops<T>(st.n, sorted + st.off, st.o, tops + (size_t)c * ASM_CHUNK, cnt,
                      words + st.off, ep, lane);
    }
}

template <typename T>
int build_segmented_t(hipStream_t s, Scratch &S, const SegBuild &g, const ft_config &cfg) {
    if (g.ntiles == 0) return FT_OK;
    // tops are indexed chunk * ASM_CHUNK: at most total + ntrees * ASM_CHUNK entries
    FT_TRY(scratch_ensure(S, g.total + (size_t)g.ntrees * ASM_CHUNK));
    if ((size_t)g.ntiles > S.lookback_words / RS_MAX_RADIX)
        return fail(FT_EINVAL, "segmented build: too many sort tiles for the scratch");
    const SegTree<T> *tr = (const SegTree<T> *)g.trees;
    const int key_bits = 3 * cfg.bits_per_dim;
    const int rbits = radix_width(S, key_bits);
    const int passes = radix_passes(key_bits, rbits);
    if (passes > RS_MAX_PASSES) return fail(FT_EINVAL, "radix sort: too many passes");
    BboxDev<T> *bb = (BboxDev<T> *)g.bbox;
    if constexpr (CoordTraits<T>::fp) {
        hipLaunchKernelGGL(seg_bbox_kernel<T>, dim3(g.ntiles), dim3(256), 0, s, tr, g.tiles,
                           (T *)g.partials);
        hipLaunchKernelGGL(seg_bbox_final_kernel<T>, dim3(g.ntrees), dim3(64), 0, s, tr,
                           (const T *)g.partials, bb);
        FT_CHECK_LAUNCH();
    }
    const size_t hw = (size_t)g.ntrees * RS_HIST_WORDS;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(hw, 256)), dim3(256), 0, s, g.hist, 0u, hw);
    hipLaunchKernelGGL(seg_encode_kernel<T>, dim3(g.ntiles), dim3(256), 0, s, tr, g.tiles,
                       cfg.bits_per_dim, cfg.sfc, (const BboxDev<T> *)bb, S.keys_in, passes, rbits,
                       g.hist, (Point4<T> *)S.points);
    uint32_t *base = g.hist + hw;
    hipLaunchKernelGGL(seg_radix_base_kernel<T>, dim3(g.ntrees), dim3(1 << rbits), 0, s, tr,
                       (const uint32_t *)g.hist, passes, rbits, base, S.tile_ctr, S.epoch);
    FT_CHECK_LAUNCH();
    // the sweeps: the last one writes the sorted keys over keys_in, which only
    // the first reads
    const uint64_t *kin = S.keys_in;
    const uint32_t *vin = nullptr;
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        uint64_t *kout = last ? S.keys_in : ((p & 1) ? S.keys_b : S.keys_a);
        uint32_t *vout = last ? S.perm : ((p & 1) ? S.vals_b : S.vals_a);
        const uint32_t *pb = base + ((size_t)p << rbits);
#define FT_SEG_SWEEP(F, BT)                                                                        \
    hipLaunchKernelGGL((onesweep_kernel<F, BT, true>), dim3(g.ntiles), dim3(RS_BLOCK), 0, s, kin, \
                       vin, kout, vout, (int)g.total, rbits * p, pb, S.lookback,                  \
                       (const uint32_t *)S.epoch, p, S.tile_ctr + p, g.tiles)
        if (rbits == 9) {
            if (p == 0) FT_SEG_SWEEP(true, 9);
            else FT_SEG_SWEEP(false, 9);
        } else {
            if (p == 0) FT_SEG_SWEEP(true, 8);
            else FT_SEG_SWEEP(false, 8);
        }
#undef FT_SEG_SWEEP
        kin = kout;
        vin = vout;
    }
    FT_CHECK_LAUNCH();
    hipLaunchKernelGGL(topo_begin_kernel, dim3(1), dim3(64), 0, s, S.topo_epoch);
    hipLaunchKernelGGL((seg_assemble_kernel<T, ASM_CHUNK>), dim3(xcd_grid(g.nchunks)),
                       dim3(ASM_CHUNK), 0, s, tr, g.chunks, g.nchunks,
                       (const uint64_t *)S.keys_in, (const uint32_t *)S.perm,
                       (const Point4<T> *)S.points, S.tops, S.chunk_tops);
    hipLaunchKernelGGL(seg_top_kernel<T>, dim3(xcd_grid((g.nchunks + 3) / 4)), dim3(256), 0, s, tr,
                       g.chunks, g.nchunks, (const uint64_t *)S.keys_in, (const int32_t *)S.tops,
                       (const uint32_t *)S.chunk_tops, S.split_words,
                       (const uint32_t *)S.topo_epoch);
    FT_CHECK_LAUNCH();
    return FT_OK;
}
template int build_segmented_t<float>(hipStream_t, Scratch &, const SegBuild &, const ft_config &);
template int build_segmented_t<double>(hipStream_t, Scratch &, const SegBuild &, const ft_config &);
template int build_segmented_t<uint32_t>(hipStream_t, Scratch &, const SegBuild &, const ft_config &);
template int build_segmented_t<uint64_t>(hipStream_t, Scratch &, const SegBuild &, const ft_config &);

}  // namespace ft

