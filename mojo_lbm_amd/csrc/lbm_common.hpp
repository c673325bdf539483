// lbm_common.hpp -- lattice tables, Float16C codec, storage traits and launch
// geometry shared by the gfx950 stream-collide kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"

namespace lbm {

// ----------------------------------------------------------------- lattices
// Velocity sets in the reference order with opposites adjacent (q odd, q+1):
// get_D2Q9 src/lbm/lattice.mojo:302-351, get_D3Q19 :221-299, get_D3Q27 :122-218.
// Tables are constexpr functions so that fully unrolled q loops fold them into
// immediates.
template <int Q_> struct Lattice;

template <> struct Lattice<9> {
    static constexpr int D = 2, Q = 9;
    __host__ __device__ static constexpr int c(int q, int d) {
        constexpr int8_t t[9][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                    {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}};
        return t[q][d];
    }
    __host__ __device__ static constexpr int wnum(int q) { return q == 0 ? 4 : 1; }
    __host__ __device__ static constexpr int wden(int q) { return q == 0 ? 9 : (q < 5 ? 9 : 36); }
};

struct Cube27 {
    __host__ __device__ static constexpr int c(int q, int d) {
        constexpr int8_t t[27][3] = {
            {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},    {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
            {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},   {1, 0, 1},   {-1, 0, -1}, {1, 0, -1},
            {-1, 0, 1},  {0, 1, 1},   {0, -1, -1}, {0, 1, -1},   {0, -1, 1},  {1, 1, 1},  {-1, -1, -1},
            {1, 1, -1},  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1},  {-1, 1, 1},  {1, -1, -1}};
        return t[q][d];
    }
    __host__ __device__ static constexpr int nnz(int q) {
        return (c(q, 0) != 0) + (c(q, 1) != 0) + (c(q, 2) != 0);
    }
};

template <> struct Lattice<19> {
    static constexpr int D = 3, Q = 19;
    __host__ __device__ static constexpr int c(int q, int d) { return Cube27::c(q, d); }
    __host__ __device__ static constexpr int wnum(int) { return 1; }
    __host__ __device__ static constexpr int wden(int q) {
        return Cube27::nnz(q) == 0 ? 3 : (Cube27::nnz(q) == 1 ? 18 : 36);
    }
};

template <> struct Lattice<27> {
    static constexpr int D = 3, Q = 27;
    __host__ __device__ static constexpr int c(int q, int d) { return Cube27::c(q, d); }
    __host__ __device__ static constexpr int wnum(int q) {
        return Cube27::nnz(q) == 0 ? 8 : (Cube27::nnz(q) == 1 ? 2 : 1);
    }
    __host__ __device__ static constexpr int wden(int q) {
        return Cube27::nnz(q) == 0 ? 27 : (Cube27::nnz(q) == 1 ? 27 : (Cube27::nnz(q) == 2 ? 54 : 216));
    }
};

// Opposite direction: the pair (q, q+1) for odd q (lattice.mojo:76-91 finds the same).
__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

// Weight in the compute type, correctly rounded (the reference writes them as literals).
template <class L, class R> __host__ __device__ __forceinline__ constexpr R weight(int q) {
    return (R)L::wnum(q) / (R)L::wden(q);
}

// ----------------------------------------------------------------- Float16C
// 1-4-11 format of Lehmann et al., PRE 106, 015308 (2022):
//   e' = 0:  |v| = m * 2^-25        e' > 0: |v| = (2048 + m) * 2^(e' - 26)
// Encoding rounds to nearest with ties away from zero and saturates to 0x7FFF
// for magnitudes that round to 2 or more (and for inf / NaN).
// Decode: the 15 magnitude bits placed at fp32 bit 12 read as an fp32 number
// equal to the value times 2^-112 (e' > 0: exponent e' - 127 instead of
// e' - 15; e' = 0: the fp32 denormal m * 2^-137 instead of m * 2^-25), so one
// exact multiply by 2^112 restores it. Needs fp32 denormal inputs, which the
// library keeps (IEEE mode, no flush). The code is sign-extended first (a signed
// 16-bit load): shifted left by 12 its sign lands on bits 27..31, and one mask
// keeps bit 31 and the magnitude.
__device__ __forceinline__ float f16c_decode_sext(int32_t hs) {
    return __uint_as_float(((uint32_t)hs << 12) & 0x87FFFFFFu) * 0x1p112f;
}
__device__ __forceinline__ float f16c_decode(uint32_t h) { return f16c_decode_sext((int16_t)(uint16_t)h); }

// Encode. The code is computed from u = |x| 2^25 (one multiply with the |.| operand
// modifier, exact for finite x below 2^103, +inf above; NaN stays NaN):
//   normal range: for a = the fp32 bits of |x|, adding half a code ulp (2^11 in fp32
//   bits), truncating 12 bits and rebiasing the exponent by 112 gives the code, which
//   rounds half away from zero; u's bits are a + (25 << 23), so
//   n = (bits(u) - 0x447FF800) >> 12 (signed, no overflow for any u >= 0) is that code,
//   negative below 2^-15 and at most the subnormal code up to 2^-14; for u = inf or NaN
//   it is above 0x7FFF;
//   subnormal range: the code rounds u half up, floor(u + 1/2), which v_cvt_rpi_i32_f32
//   computes without rounding the sum (0 for NaN, saturated for u >= 2^31); clamped to
//   2048, the first normal code, it never exceeds the normal code above 2^-14.
// The code is the larger of the two; v_cvt_pk_i16_i32 saturates it to 0x7FFF (which
// saturates magnitudes of 2 - 2^-12 and above, infinities and NaN) and places it in the
// upper half, where one v_and_or_b32 adds the sign. Eight instructions; the result's upper
// half is the code (stored with a d16_hi store). Checked against the format definition on
// all 2^32 inputs.
__device__ __forceinline__ uint32_t f16c_encode_hi(float x) {
    const float u = __builtin_fabsf(x) * 0x1p25f;
    int32_t sub;
    __asm__("v_cvt_rpi_i32_f32 %0, %1" : "=v"(sub) : "v"(u));
    const int32_t n = (int32_t)(__float_as_uint(u) - 0x447FF800u) >> 12;
    const int32_t code = max(n, min(sub, 2048));
    uint32_t pk;  // upper half: min(code, 0x7FFF) (code >= 0)
    __asm__("v_cvt_pk_i16_i32 %0, %1, %2" : "=v"(pk) : "v"(code), "v"(code));
    return (__float_as_uint(x) & 0x80000000u) | pk;
}
__device__ __forceinline__ uint32_t f16c_encode(float x) { return f16c_encode_hi(x) >> 16; }

// ----------------------------------------------------------------- storage
// Populations are stored as S::T and computed in S::R (FP32 for F32 and F16C).
struct StoreF32 {
    using T = float;
    using R = float;
    static constexpr int id = LBM_STORE_F32;
    __device__ __forceinline__ static R ld(const T* p) { return *p; }
    __device__ __forceinline__ static void st(T* p, R v) { __builtin_nontemporal_store(v, p); }
};
struct StoreF16C {
    using T = uint16_t;
    using R = float;
    static constexpr int id = LBM_STORE_F16C;
    __device__ __forceinline__ static R ld(const T* p) { return f16c_decode(*p); }
    __device__ __forceinline__ static void st(T* p, R v) {
        __builtin_nontemporal_store((uint16_t)(f16c_encode_hi(v) >> 16), p);  // global_store_short_d16_hi
    }
};
struct StoreF64 {
    using T = double;
    using R = double;
    static constexpr int id = LBM_STORE_F64;
    __device__ __forceinline__ static R ld(const T* p) { return *p; }
    __device__ __forceinline__ static void st(T* p, R v) { __builtin_nontemporal_store(v, p); }
};

// ----------------------------------------------------------------- geometry
// Launch geometry of one run. Tile dimensions are powers of two; one workgroup
// processes one tile, one lane one node, lane id = local column-major index.
struct Geo {
    int nx, ny, nz;
    int txb, tyb, tzb;    // log2 of the tile dimensions
    int Tx, Ty, Tz;       // tiles per axis
    int tile_cells;       // tx * ty * tz
    int n_tiles;
    int xcd_chunk;        // tiles per XCD when the XCD-contiguous remap is on, else 0
    int wzb;              // log2 of the z layers one step workgroup covers (tzb: a whole tile)
    uint32_t hx_magic;    // floor(2^32 / (tx + 2)) + 1: n / (tx + 2) == umulhi(n, hx_magic) for n < 2^16
    uint32_t hxy_magic;   // the same for (tx + 2) * (ty + 2)
};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X. With the
// remap, XCD k walks a contiguous run of tiles, so the x- and y-neighbour tiles
// whose halo cache lines a tile shares are resident in the same L2. With 2^s
// workgroups per tile the result is the work unit tile * 2^s + part, and the
// parts of one tile stay on one XCD.
__device__ __forceinline__ int tile_of_block(const Geo& g, int s = 0) {
    const int b = blockIdx.x;
    return g.xcd_chunk ? (b & 7) * (g.xcd_chunk << s) + (b >> 3) : b;
}

}  // namespace lbm
