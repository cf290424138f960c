// Written by gen_dpp_fma.py.
//
// Row-broadcast FP64 FMAs for gfx950 (CDNA4).  The 64 lanes of a wavefront form four rows of
// 16; a DP-ALU DPP operand with row_newbcast:k reads lane k of the lane's own row.  A row
// whose lanes hold the E values x[0..E-1] (lanes 0..15 of register ka, then lanes 0..E-17 of
// kb when E > 16) feeds every lane's E FMAs with no scalar or LDS traffic.  Defined for the
// odd row lengths E = 5 .. 17 of the far-field subgrids.
//
// Each function is one asm statement: it starts with s_nop 1, the two wait states a DPP read
// needs after a VALU write of its source (the compiler does not check hazards inside inline
// asm), and every lane of the wavefront must be active.
#pragma once

namespace pdmk::gpu {

// acc[k] += x[k] * c, k = 0 .. E-1
template <int E> __device__ __forceinline__ void dpp_row_axpy(double (&acc)[E], double ka, double kb, double c);
// s[k & 1] += x[k] * v[k], k = 0 .. E-1
template <int E> __device__ __forceinline__ void dpp_row_dot(double &s0, double &s1, double ka, double kb, const double (&v)[E]);

template <> __device__ __forceinline__ void dpp_row_axpy<5>(double (&acc)[5], double ka, double kb, double c) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_axpy<7>(double (&acc)[7], double ka, double kb, double c) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a5], %[ka], %[c] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a6], %[ka], %[c] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4]), [a5] "+v"(acc[5]), [a6] "+v"(acc[6])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_axpy<9>(double (&acc)[9], double ka, double kb, double c) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a5], %[ka], %[c] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a6], %[ka], %[c] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a7], %[ka], %[c] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a8], %[ka], %[c] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4]), [a5] "+v"(acc[5]), [a6] "+v"(acc[6]), [a7] "+v"(acc[7]), [a8] "+v"(acc[8])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_axpy<11>(double (&acc)[11], double ka, double kb, double c) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a5], %[ka], %[c] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a6], %[ka], %[c] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a7], %[ka], %[c] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a8], %[ka], %[c] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a9], %[ka], %[c] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a10], %[ka], %[c] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4]), [a5] "+v"(acc[5]), [a6] "+v"(acc[6]), [a7] "+v"(acc[7]), [a8] "+v"(acc[8]), [a9] "+v"(acc[9]), [a10] "+v"(acc[10])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_axpy<13>(double (&acc)[13], double ka, double kb, double c) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a5], %[ka], %[c] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a6], %[ka], %[c] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a7], %[ka], %[c] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a8], %[ka], %[c] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a9], %[ka], %[c] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a10], %[ka], %[c] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a11], %[ka], %[c] row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a12], %[ka], %[c] row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4]), [a5] "+v"(acc[5]), [a6] "+v"(acc[6]), [a7] "+v"(acc[7]), [a8] "+v"(acc[8]), [a9] "+v"(acc[9]), [a10] "+v"(acc[10]), [a11] "+v"(acc[11]), [a12] "+v"(acc[12])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_axpy<15>(double (&acc)[15], double ka, double kb, double c) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a5], %[ka], %[c] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a6], %[ka], %[c] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a7], %[ka], %[c] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a8], %[ka], %[c] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a9], %[ka], %[c] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a10], %[ka], %[c] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a11], %[ka], %[c] row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a12], %[ka], %[c] row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a13], %[ka], %[c] row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a14], %[ka], %[c] row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4]), [a5] "+v"(acc[5]), [a6] "+v"(acc[6]), [a7] "+v"(acc[7]), [a8] "+v"(acc[8]), [a9] "+v"(acc[9]), [a10] "+v"(acc[10]), [a11] "+v"(acc[11]), [a12] "+v"(acc[12]), [a13] "+v"(acc[13]), [a14] "+v"(acc[14])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_axpy<17>(double (&acc)[17], double ka, double kb, double c) {
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[a0], %[ka], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a1], %[ka], %[c] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a2], %[ka], %[c] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a3], %[ka], %[c] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a4], %[ka], %[c] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a5], %[ka], %[c] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a6], %[ka], %[c] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a7], %[ka], %[c] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a8], %[ka], %[c] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a9], %[ka], %[c] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a10], %[ka], %[c] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a11], %[ka], %[c] row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a12], %[ka], %[c] row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a13], %[ka], %[c] row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a14], %[ka], %[c] row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a15], %[ka], %[c] row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[a16], %[kb], %[c] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        : [a0] "+v"(acc[0]), [a1] "+v"(acc[1]), [a2] "+v"(acc[2]), [a3] "+v"(acc[3]), [a4] "+v"(acc[4]), [a5] "+v"(acc[5]), [a6] "+v"(acc[6]), [a7] "+v"(acc[7]), [a8] "+v"(acc[8]), [a9] "+v"(acc[9]), [a10] "+v"(acc[10]), [a11] "+v"(acc[11]), [a12] "+v"(acc[12]), [a13] "+v"(acc[13]), [a14] "+v"(acc[14]), [a15] "+v"(acc[15]), [a16] "+v"(acc[16])
        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));
}
template <> __device__ __forceinline__ void dpp_row_dot<5>(double &s0, double &s1, double ka, double kb, const double (&v)[5]) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]));
}
template <> __device__ __forceinline__ void dpp_row_dot<7>(double &s0, double &s1, double ka, double kb, const double (&v)[7]) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v5] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v6] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]), [v5] "v"(v[5]), [v6] "v"(v[6]));
}
template <> __device__ __forceinline__ void dpp_row_dot<9>(double &s0, double &s1, double ka, double kb, const double (&v)[9]) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v5] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v6] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v7] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v8] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]), [v5] "v"(v[5]), [v6] "v"(v[6]), [v7] "v"(v[7]), [v8] "v"(v[8]));
}
template <> __device__ __forceinline__ void dpp_row_dot<11>(double &s0, double &s1, double ka, double kb, const double (&v)[11]) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v5] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v6] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v7] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v8] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v9] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v10] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]), [v5] "v"(v[5]), [v6] "v"(v[6]), [v7] "v"(v[7]), [v8] "v"(v[8]), [v9] "v"(v[9]), [v10] "v"(v[10]));
}
template <> __device__ __forceinline__ void dpp_row_dot<13>(double &s0, double &s1, double ka, double kb, const double (&v)[13]) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v5] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v6] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v7] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v8] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v9] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v10] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v11] row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v12] row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]), [v5] "v"(v[5]), [v6] "v"(v[6]), [v7] "v"(v[7]), [v8] "v"(v[8]), [v9] "v"(v[9]), [v10] "v"(v[10]), [v11] "v"(v[11]), [v12] "v"(v[12]));
}
template <> __device__ __forceinline__ void dpp_row_dot<15>(double &s0, double &s1, double ka, double kb, const double (&v)[15]) {
    (void)kb;
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v5] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v6] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v7] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v8] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v9] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v10] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v11] row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v12] row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v13] row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v14] row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]), [v5] "v"(v[5]), [v6] "v"(v[6]), [v7] "v"(v[7]), [v8] "v"(v[8]), [v9] "v"(v[9]), [v10] "v"(v[10]), [v11] "v"(v[11]), [v12] "v"(v[12]), [v13] "v"(v[13]), [v14] "v"(v[14]));
}
template <> __device__ __forceinline__ void dpp_row_dot<17>(double &s0, double &s1, double ka, double kb, const double (&v)[17]) {
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v0] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v1] row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v2] row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v3] row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v4] row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v5] row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v6] row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v7] row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v8] row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v9] row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v10] row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v11] row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v12] row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v13] row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[ka], %[v14] row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s1], %[ka], %[v15] row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %[s0], %[kb], %[v16] row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        : [s0] "+v"(s0), [s1] "+v"(s1)
        : [ka] "v"(ka), [kb] "v"(kb), [v0] "v"(v[0]), [v1] "v"(v[1]), [v2] "v"(v[2]), [v3] "v"(v[3]), [v4] "v"(v[4]), [v5] "v"(v[5]), [v6] "v"(v[6]), [v7] "v"(v[7]), [v8] "v"(v[8]), [v9] "v"(v[9]), [v10] "v"(v[10]), [v11] "v"(v[11]), [v12] "v"(v[12]), [v13] "v"(v[13]), [v14] "v"(v[14]), [v15] "v"(v[15]), [v16] "v"(v[16]));
}

}  // namespace pdmk::gpu
