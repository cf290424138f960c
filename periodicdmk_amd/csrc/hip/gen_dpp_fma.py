"""Writes dpp_fma.hpp (row-broadcast FP64 FMA blocks for the far-field NUFFT kernels).

Run from anywhere: python gen_dpp_fma.py
"""
import os

lines=[]
A=lines.append
A('''// Written by gen_dpp_fma.py.
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
''')
def fm(dst, src1, k):
    src = 'ka' if k < 16 else 'kb'
    return '"v_fmac_f64_dpp %%[%s], %%[%s], %%[%s] row_newbcast:%d row_mask:0xf bank_mask:0xf\\n"' % (dst, src, src1, k % 16)
ES=range(5,18,2)
for E in ES:
    outs=', '.join('[a%d] "+v"(acc[%d])'%(k,k) for k in range(E))
    A('template <> __device__ __forceinline__ void dpp_row_axpy<%d>(double (&acc)[%d], double ka, double kb, double c) {' % (E,E))
    if E <= 16: A('    (void)kb;')
    A('    asm("s_nop 1\\n"')
    for k in range(E): A('        ' + fm('a%d'%k,'c',k))
    A('        : ' + outs)
    A('        : [ka] "v"(ka), [kb] "v"(kb), [c] "v"(c));')
    A('}')
for E in ES:
    ins=', '.join('[v%d] "v"(v[%d])'%(k,k) for k in range(E))
    A('template <> __device__ __forceinline__ void dpp_row_dot<%d>(double &s0, double &s1, double ka, double kb, const double (&v)[%d]) {' % (E,E))
    if E <= 16: A('    (void)kb;')
    A('    asm("s_nop 1\\n"')
    for k in range(E): A('        ' + fm('s%d'%(k&1),'v%d'%k,k))
    A('        : [s0] "+v"(s0), [s1] "+v"(s1)')
    A('        : [ka] "v"(ka), [kb] "v"(kb), ' + ins + ');')
    A('}')
A('')
A('}  // namespace pdmk::gpu')
open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'dpp_fma.hpp'),'w').write('\n'.join(lines)+'\n')
