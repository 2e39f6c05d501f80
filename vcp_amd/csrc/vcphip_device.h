// Device-side helpers shared by the libvcphip kernels (gfx950 / CDNA4).
//
// Directed rounding.  Every kernel that needs directed rounding runs its
// whole body in round-toward-+inf: it writes the MODE register once at entry
// (FP_ROUND bits [3:0]: f32 and f64/f16 both := 1, toward +inf) and obtains
// round-toward-(-inf) through RD(x) = -RU(-x), which is exact.  The write is
// an inline-asm s_setreg so that the compiler's mode tracking (which assumes
// the default round-to-nearest for ordinary FP instructions and would
// otherwise re-insert a switch back to nearest before them) stays out of the
// way; all FP work in such kernels depends on memory loads issued after the
// asm (it carries a memory clobber), so nothing can be scheduled above it.
// Sources are built with -ffp-contract=off: every rounding is explicit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VCPHIP_DEVINL __device__ __forceinline__

// LDS byte address of a __shared__ object.
VCPHIP_DEVINL uint32_t vcphip_lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
// One wave-wide 16-byte-per-lane load straight into LDS
// (global_load_lds_dwordx4): lane l's 16 bytes, from its own address g, land
// at lds_base + 16 l (lds_base must be wave-uniform).  Neither the load nor
// its completion is visible to the compiler: callers wait with an explicit
// s_waitcnt vmcnt before a barrier.  m0 is listed as clobbered; hipcc warns
// that a reserved register in a clobber list is not preserved for code that
// keeps a value in it.  The kernels using this helper keep nothing in m0:
// in their gfx950 assembly m0 is written only by this sequence.
VCPHIP_DEVINL void vcphip_lds_dma16(const void* g, uint32_t lds_base) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(lds_base) : "memory", "m0");
}

// A wave-uniform double as an opaque register value (v_readfirstlane of its
// halves), so that selects between such values stay selects of values.
VCPHIP_DEVINL double vcphip_uniform(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Small integer helpers of the residue kernels.
// a b + c with a and b taken as signed 24-bit values (v_mad_i32_i24); b is
// an SGPR operand and must be wave-uniform (the callers pass modulus constants)
VCPHIP_DEVINL unsigned vcphip_mad_i24(unsigned a, int b, unsigned c) {
  unsigned d;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
  return d;
}
// (a ^ b) + c (v_xad_u32): with b = -1, c = 1 the negation of a
VCPHIP_DEVINL unsigned vcphip_xad(unsigned a, unsigned b, unsigned c) {
  unsigned d;
  asm("v_xad_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
// floor(y / m) = (8 y magic) >> 32 for magic = ceil(2^29 / m), y8 = 8 y <
// 2^24 and m <= 256: with magic = 2^29 / m + d, d < 1, the product is
// 8 y / m + 8 y d / 2^32 < y / m + 2^-8, and frac(y / m) <= 1 - 1/m leaves
// room for that for m < 256 (d = 0 for m = 256)  (v_mul_hi_u32_u24)
VCPHIP_DEVINL unsigned vcphip_quot24(unsigned y8, unsigned magic) {
  return (unsigned)(((unsigned long long)(y8 & 0xFFFFFFu) * (unsigned long long)(magic & 0xFFFFFFu)) >> 32);
}
// the low bytes of r[0..3] as one word, r[0] lowest
VCPHIP_DEVINL unsigned vcphip_pack4(const unsigned (&r)[4]) {
  const unsigned lo = __builtin_amdgcn_perm(r[1], r[0], 0x0c0c0400u);   // bytes r0, r1, 0, 0
  const unsigned hi = __builtin_amdgcn_perm(r[3], r[2], 0x04000c0cu);   // bytes 0, 0, r2, r3
  return lo | hi;
}

// Raw buffer resource (stride 0, gfx9 word 3) over `bytes` bytes from a
// wave-uniform base: loads through it take a 32-bit lane offset in bytes and
// need no per-lane 64-bit address arithmetic; offsets at or past `bytes`
// read 0.
VCPHIP_DEVINL __amdgpu_buffer_rsrc_t vcphip_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}

VCPHIP_DEVINL void vcphip_mode_round_up() {
  asm volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 4), 5" ::: "memory");
}

// 2^e as a double for e in [-1022, 1023] (bit construction, exact)
VCPHIP_DEVINL double vcphip_pow2(int e) {
  return __longlong_as_double((long long)(e + 1023) << 52);
}

// x * 2^e under the current rounding mode; intermediate factors stay normal
// so that only an underflowing final product rounds.
VCPHIP_DEVINL double vcphip_mulpow2(double x, int e) {
  while (e > 1000) { x = x * vcphip_pow2(1000); e -= 1000; }
  while (e < -1000) { x = x * vcphip_pow2(-1000); e += 1000; }
  return x * vcphip_pow2(e);
}

// frexp exponent of x > 0: 2^(e-1) <= x < 2^e
VCPHIP_DEVINL int vcphip_fexp(double x) {
  return __builtin_amdgcn_frexp_exp(x);
}

// midpoint-radius form of [lo, hi] in round-up mode: [m-r, m+r] contains
// [lo, hi]; point intervals give (lo, 0).  Same rule as the oracle's
// restatement of kv's midrad.
VCPHIP_DEVINL void vcphip_midrad_ru(double lo, double hi, double& m, double& r) {
  if (lo == hi) {
    m = lo;
    r = 0.0;
  } else {
    const double h1 = 0.5 * lo;
    const double h2 = 0.5 * hi;
    m = h1 + h2;
    const double r1 = hi - m;
    const double r2 = m - lo;
    r = r1 > r2 ? r1 : r2;
  }
}

// XCD-aware block remap: workgroups are dispatched round-robin over the 8
// XCDs of an MI355X; give each XCD a contiguous range of tile ids so that
// concurrently running tiles on one XCD share A rows / B columns in its L2.
VCPHIP_DEVINL int vcphip_xcd_remap(int bid, int ntiles) {
  const int q = ntiles >> 3, r = ntiles & 7;
  const int x = bid & 7, local = bid >> 3;
  const int start = x * q + (x < r ? x : r);
  return start + local;
}

// grouped tile order: GH tile-rows per group, columns inside a group
VCPHIP_DEVINL void vcphip_tile_coords(int tile, int ntm, int ntn, int GH, int& tm, int& tn) {
  const int per_group = GH * ntn;
  const int group = tile / per_group;
  const int idx = tile - group * per_group;
  int gh = ntm - group * GH;
  gh = gh < GH ? gh : GH;
  tm = group * GH + idx % gh;
  tn = idx / gh;
}

// byte offset of (row, k) in an int8 operand plane of the Ozaki GEMMs: 1 KB
// blocks of 16 rows x 64 k (row-major inside), ordered by row group, then k
VCPHIP_DEVINL long vcphip_plane_off(long row, int k, int kpad) {
  return ((row >> 4) * (long)(kpad >> 6) + (k >> 6)) * 1024 + (row & 15) * 64 + (k & 63);
}
