// fgl_draws.h -- where a draw kernel finds its draw and its block index.
//
// A draw's kernels are written once, as bodies taking the draw's arguments
// and a grid view G, and launched two ways:
//   OneDraw    the kernel's own arguments are the draw; the block is
//              blockIdx.x of gridDim.x.
//   ManyDraws  one launch covers many draws (the views of a batch): launch
//              block b belongs to draw v = view_of[b] as its block
//              b - first[v] of first[v + 1] - first[v].  The draws' arguments
//              sit in a device table that no kernel writes, read through the
//              constant address space so that, like kernel arguments, they
//              are fetched with scalar loads.
#pragma once

#include "fgl_internal.h"

namespace fgl {

typedef const __attribute__((address_space(4))) DrawArgs ConstDrawArgs;
typedef const __attribute__((address_space(4))) uint32_t ConstU32;

// The kernel's arguments read again from the kernarg segment.  The empty
// asm makes the pointer opaque, so the loads through it stay at their uses
// (scalar loads each pass of a loop) instead of being hoisted in front of
// the loop into scalar registers that the loop then spills.
FGL_D const DrawArgs &args_reload() {
    auto p = (ConstDrawArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const DrawArgs *)p;
}

struct OneDraw {
    FGL_D uint32_t bid() const { return blockIdx.x; }
    FGL_D uint32_t nblk() const { return gridDim.x; }
    FGL_D const DrawArgs &args() const { return args_reload(); }
};

struct ManyDraws {
    uint32_t b, n;
    ConstDrawArgs *p;
    FGL_D uint32_t bid() const { return b; }
    FGL_D uint32_t nblk() const { return n; }
    FGL_D const DrawArgs &args() const {
        ConstDrawArgs *q = p;
        asm volatile("" : "+s"(q));
        return *(const DrawArgs *)q;
    }
};

FGL_D ManyDraws many_draws(const MultiArgs &m) {
    const uint32_t v = ((ConstU32 *)m.view_of)[blockIdx.x];
    const uint32_t f0 = ((ConstU32 *)m.first)[v], f1 = ((ConstU32 *)m.first)[v + 1];
    return ManyDraws{blockIdx.x - f0, f1 - f0, (ConstDrawArgs *)m.draws + v};
}

FGL_D const DrawArgs &draw_of(const ManyDraws &g) { return *(const DrawArgs *)g.p; }

}  // namespace fgl

// The two kernels of a draw stage whose body is BODY(const DrawArgs &, const
// G &): NAME (one draw, its arguments by value) and NAME_m (many draws).
#define FGL_DRAW_KERNELS(NAME, BODY, BOUNDS)                                        \
    __global__ BOUNDS void NAME(DrawArgs a) { BODY(a, OneDraw{}); }                 \
    __global__ BOUNDS void NAME##_m(MultiArgs m) {                                  \
        const ManyDraws g = many_draws(m);                                          \
        BODY(draw_of(g), g);                                                        \
    }
// The same for a radix stage, whose body also takes the pass.
#define FGL_DRAW_KERNELS_PASS(NAME, BODY, BOUNDS)                                   \
    __global__ BOUNDS void NAME(DrawArgs a, int pass) { BODY(a, OneDraw{}, pass); } \
    __global__ BOUNDS void NAME##_m(MultiArgs m) {                                  \
        const ManyDraws g = many_draws(m);                                          \
        BODY(draw_of(g), g, m.pass);                                                \
    }
