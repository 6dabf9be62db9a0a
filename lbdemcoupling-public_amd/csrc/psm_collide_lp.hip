/*
 * FP32-deviation and FP16-storage instantiations of the fused PSM
 * collide + stream kernel.  This unit is compiled with FMA contraction
 * (-ffp-contract=fast) and uses the hardware reciprocal: these variants are
 * checked against FP64 by error bounds, not bit for bit.
 */
#include "psm_collide_kernel.h"

namespace psm {

template <int PREC>
static void launch_lp(lbdem_psm* h, const CollideArgs& a, const void* vin, void* vout) {
    typedef typename Store<PREC>::S S;
    const S* fin = (const S*)vin;
    S* fout = (S*)vout;
    const int wgz = collide_wg(a.L.zc);
    dim3 grid((unsigned)((int64_t)a.nx * a.ny * a.L.nzc * (a.L.zc / wgz))), block((unsigned)wgz);
    if (!h->cs_lds_ready) {
        const int dev = h->desc.device, cap = collide_wave_cap(PREC);
        h->cs_lds[1][1] = collide_lds_for(k_collide_stream<PREC, true, true>, dev, wgz, cap);
        h->cs_lds[1][0] = collide_lds_for(k_collide_stream<PREC, true, false>, dev, wgz, cap);
        h->cs_lds[0][1] = collide_lds_for(k_collide_stream<PREC, false, true>, dev, wgz, cap);
        h->cs_lds[0][0] = collide_lds_for(k_collide_stream<PREC, false, false>, dev, wgz, cap);
        h->cs_lds_ready = true;
    }
    const unsigned lds = h->cs_lds[h->has_force ? 1 : 0][a.pz ? 1 : 0];
    if (h->has_force) {
        if (a.pz) hipLaunchKernelGGL((k_collide_stream<PREC, true, true>), grid, block, lds, h->stream, fin, fout, a);
        else hipLaunchKernelGGL((k_collide_stream<PREC, true, false>), grid, block, lds, h->stream, fin, fout, a);
    } else {
        if (a.pz) hipLaunchKernelGGL((k_collide_stream<PREC, false, true>), grid, block, lds, h->stream, fin, fout, a);
        else hipLaunchKernelGGL((k_collide_stream<PREC, false, false>), grid, block, lds, h->stream, fin, fout, a);
    }
}

template <>
void launch_cs_prec<LBDEM_PSM_FP32_DEV>(lbdem_psm* h, const CollideArgs& a, const void* fin, void* fout) {
    launch_lp<LBDEM_PSM_FP32_DEV>(h, a, fin, fout);
}
template <>
void launch_cs_prec<LBDEM_PSM_FP16_STORE>(lbdem_psm* h, const CollideArgs& a, const void* fin, void* fout) {
    launch_lp<LBDEM_PSM_FP16_STORE>(h, a, fin, fout);
}

}  // namespace psm
