// sw2_batch.hip -- the batch kernels of the mixed-precision step on one tile
// shape, SW2_TILE_X x SW2_TILE_Y columns x rows, and their launcher
// SW2_BATCH_PHASES (sw2_batch.h).  The Makefile compiles this file once per
// shape (32 x 16 with the persistent kernels, 64 x 8, 64 x 16), each into a
// namespace of its own (SW2_SHAPE_NS), so the shapes compile in parallel and
// apart from the single-run kernels of sw2_device.hip.  The kernels are the
// tile bodies of sw2_mixed.inc, the ones the single-run mixed kernels run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "sw2_batch.h"
#include "sw2_internal.h"
#include "sw2_xdiv.h"

#if !defined(SW2_TILE_X) || !defined(SW2_TILE_Y) || !defined(SW2_SHAPE_NS) || !defined(SW2_BATCH_PHASES)
#error "sw2_batch.hip is compiled with the tile shape, its namespace and its launcher's name defined"
#endif

namespace {
#include "sw2_base.inc"
}  // namespace

namespace SW2_SHAPE_NS {
#include "sw2_tile.inc"
}  // namespace SW2_SHAPE_NS
using namespace SW2_SHAPE_NS;
#include "sw2_mixed_common.inc"
#define SW2_BATCH_KERNELS
namespace SW2_SHAPE_NS {
#include "sw2_mixed.inc"

template <int SCH, int HV, bool TR>
static void phases(const BMember *mem, const BState *st, const BHdr &hdr, int ge, int gu, int pe, int pu, hipStream_t s) {
#ifdef SW2_PERSISTENT
    if (pe) kbp_eta<SCH, HV, TR><<<ge, NTE, 0, s>>>(mem, st, hdr);
    else kb_eta<SCH, HV, TR><<<ge, NTE, 0, s>>>(mem, st, hdr);
    if (pu) kbp_uv<SCH, HV, TR><<<gu, NTU, 0, s>>>(mem, st, hdr);
    else kb_uv<SCH, HV, TR><<<gu, NTU, 0, s>>>(mem, st, hdr);
#else
    (void)pe;
    (void)pu;
    kb_eta<SCH, HV, TR><<<ge, NTE, 0, s>>>(mem, st, hdr);
    kb_uv<SCH, HV, TR><<<gu, NTU, 0, s>>>(mem, st, hdr);
#endif
}

template <int SCH>
static void phases_sch(const BMember *mem, const BState *st, const BHdr &hdr, int hv, int tr, int ge, int gu, int pe,
                       int pu, hipStream_t s) {
    if (tr) {
        if (hv) phases<SCH, 2, true>(mem, st, hdr, ge, gu, pe, pu, s);
        else phases<SCH, 0, true>(mem, st, hdr, ge, gu, pe, pu, s);
    } else {
        if (hv) phases<SCH, 2, false>(mem, st, hdr, ge, gu, pe, pu, s);
        else phases<SCH, 0, false>(mem, st, hdr, ge, gu, pe, pu, s);
    }
}
}  // namespace SW2_SHAPE_NS

SW2_BATCH_LAUNCHER(SW2_BATCH_PHASES) {
    const BMember *m = (const BMember *)mem;
    const BState *b = (const BState *)st;
    const BHdr &h = *(const BHdr *)hdr;  // the kernels take it by value
    switch (sch) {
        case 0: SW2_SHAPE_NS::phases_sch<0>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 1: SW2_SHAPE_NS::phases_sch<1>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 2: SW2_SHAPE_NS::phases_sch<2>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 3: SW2_SHAPE_NS::phases_sch<3>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 4: SW2_SHAPE_NS::phases_sch<4>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 5: SW2_SHAPE_NS::phases_sch<5>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 6: SW2_SHAPE_NS::phases_sch<6>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 7: SW2_SHAPE_NS::phases_sch<7>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        case 8: SW2_SHAPE_NS::phases_sch<8>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
        default: SW2_SHAPE_NS::phases_sch<9>(m, b, h, hv, tr, ge, gu, pe, pu, s); break;
    }
}
