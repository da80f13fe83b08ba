// sw2_device.hip -- GPU time step of SW2 for gfx950 (MI355X), FP64 form.
//
// One device session holds a world's fields in HBM with SW2's sppde layout
// (halo of 2, x fastest) and advances it with the per-step order of loop()
// (core.c:134-338).  Each phase is a kernel over the interior with one lane
// per cell, longitude-major so that a wavefront reads 64 consecutive cells of
// a row.  Halo updates reproduce the single-map rules of
// sppde/sppde_hupdate.c:274-363 (periodic images in x and in y, corners
// included) and do_channel (calc.c:600-649) rewrites interior columns only,
// so the halo state -- and with it every value -- is the one SW2's CPU path
// computes on one rank.
//
// Arithmetic: FP64, compiled with -ffp-contract=off so no multiply-add is
// fused; every expression keeps SW2's association; divisions are IEEE
// correctly rounded; transcendental row values come from the host libm.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <utility>

#include "sw2_batch.h"
#include "sw2_internal.h"
#include "sw2_xdiv.h"

#define HIPCHECK(x)                                                                     \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            sw2_set_error("HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                  \
        }                                                                               \
    } while (0)

namespace {
#include "sw2_base.inc"

struct Ptrs {
    double *eta, *u, *v, *wx, *wy, *tr;
};

#define CELL()                                                   \
    const int i = 1 + blockIdx.x * BX + threadIdx.x;             \
    const int j = 1 + blockIdx.y * BY + threadIdx.y;             \
    if (fl->fail_step) return;                                   \
    if (i > g.nx || j > g.ny) return;

#define UW(i, j) (p.u[IX(g, i, j)] + p.wx[IX(g, i, j)])
#define VW(i, j) (p.v[IX(g, i, j)] + p.wy[IX(g, i, j)])

// ------------------------------------------------ CFL guard (calc.c:701-738)
__global__ __launch_bounds__(BX *BY) void k_cfl(Grid g, Ptrs p, Flags *fl, int slot) {
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.y * BY + threadIdx.y;
    if (fl->fail_step) return;
    double mx = 0, my = 0;
    if (i <= g.nx && j <= g.ny) {
        double ax = ddiv(fabs(UW(i, j)), TB(g, SW2_T_DXE, j));
        double ay = ddiv(fabs(VW(i, j)), TB(g, SW2_T_DYN, j));
        if (ax > mx) mx = ax;  // NaN never wins a comparison, as in SW2
        if (ay > my) my = ay;
    }
    for (int o = 32; o > 0; o >>= 1) {
        double ox = __shfl_xor(mx, o), oy = __shfl_xor(my, o);
        mx = ox > mx ? ox : mx;
        my = oy > my ? oy : my;
    }
    __shared__ double sx[BY], sy[BY];
    if (threadIdx.x == 0) {
        sx[threadIdx.y] = mx;
        sy[threadIdx.y] = my;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        for (int q = 1; q < BY; q++) {
            mx = sx[q] > mx ? sx[q] : mx;
            my = sy[q] > my ? sy[q] : my;
        }
        // non-negative doubles order like their bit patterns
        atomicMax(&fl->cfl[2 * slot], (unsigned long long)__double_as_longlong(mx));
        atomicMax(&fl->cfl[2 * slot + 1], (unsigned long long)__double_as_longlong(my));
    }
}

__global__ void k_cfl_fin(Flags *fl, int slot, int n, double courant, double Dt, double rest_x, double rest_y) {
    if (fl->fail_step) return;
    double mx = __longlong_as_double((long long)fl->cfl[2 * slot]);
    double my = __longlong_as_double((long long)fl->cfl[2 * slot + 1]);
    if (mx + my == 0) {  // flow at rest: gravity-wave speeds (calc.c:721-729)
        mx = rest_x;
        my = rest_y;
    }
    double dtcfl = ddiv(courant, (mx + my));
    fl->last_dtcfl = dtcfl;
    fl->cfl[2 * (slot ^ 1)] = 0;
    fl->cfl[2 * (slot ^ 1) + 1] = 0;
    if (Dt > dtcfl) {
        fl->fail_step = n;
        fl->fail_kind = 1;
        fl->fail_val = dtcfl;
    }
}

// --------------------------------------- conservative flux body (BODYH, BODYTRACER)
// B(i,j) is H = (eta+depth)-hB for h (hB is identically zero) or the tracer.
template <int SCH, bool TRACER>
__global__ __launch_bounds__(BX *BY) void k_flux(Grid g, Ptrs p, Flags *fl, const double *B0, double depth, double Dt,
                                                   double sig2, double *out, int gated) {
    if (gated && !fl->nonfinite) return;
    CELL();
    auto B = [&](int a, int b) -> double {
        if constexpr (TRACER) return B0[IX(g, a, b)];
        else return ((B0[IX(g, a, b)]) + (depth));
    };
    double dxe = TB(g, SW2_T_DXE, j), dyn = TB(g, SW2_T_DYN, j);
    double uwm = UW(i - 1, j), uw0 = UW(i, j), vwm = VW(i, j - 1), vw0 = VW(i, j);
    double CwP = ddiv(0.5 * (uwm + fabs(uwm)) * Dt, dxe);
    double CwM = ddiv(0.5 * (uwm - fabs(uwm)) * Dt, dxe);
    double CeP = ddiv(0.5 * (uw0 + fabs(uw0)) * Dt, dxe);
    double CeM = ddiv(0.5 * (uw0 - fabs(uw0)) * Dt, dxe);
    double CsP = ddiv(0.5 * (vwm + fabs(vwm)) * Dt, dyn);
    double CsM = ddiv(0.5 * (vwm - fabs(vwm)) * Dt, dyn);
    double CnP = ddiv(0.5 * (vw0 + fabs(vw0)) * Dt, dyn);
    double CnM = ddiv(0.5 * (vw0 - fabs(vw0)) * Dt, dyn);
    double bxm2 = B(i - 2, j), bxm1 = B(i - 1, j), b00 = B(i, j), bxp1 = B(i + 1, j), bxp2 = B(i + 2, j);
    double bym2 = B(i, j - 2), bym1 = B(i, j - 1), byp1 = B(i, j + 1), byp2 = B(i, j + 2);
    double rxPprev = r2rP(bxm2, bxm1, b00);
    double rxMprev = r2rM(bxm1, b00, bxp1);
    double rxP = r2rP(bxm1, b00, bxp1);
    double rxM = r2rM(b00, bxp1, bxp2);
    double ryPprev = r2rP(bym2, bym1, b00);
    double ryMprev = r2rM(bym1, b00, byp1);
    double ryP = r2rP(bym1, b00, byp1);
    double ryM = r2rM(b00, byp1, byp2);
    double BwP = b2bp<SCH>(bxm1, b00, rxPprev, CwP);
    double BwM = b2bm<SCH>(bxm1, b00, rxMprev, CwM);
    double BeP = b2bp<SCH>(b00, bxp1, rxP, CeP);
    double BeM = b2bm<SCH>(b00, bxp1, rxM, CeM);
    double BsP = b2bp<SCH>(bym1, b00, ryPprev, CsP);
    double BsM = b2bm<SCH>(bym1, b00, ryMprev, CsM);
    double BnP = b2bp<SCH>(b00, byp1, ryP, CnP);
    double BnM = b2bm<SCH>(b00, byp1, ryM, CnM);
    double d = CwP * BwP + CwM * BwM - CeP * BeP - CeM * BeM + CsP * BsP + CsM * BsM - CnP * BnP - CnM * BnM;
    if constexpr (TRACER) {
        double t0 = B0[IX(g, i, j)];
        double pv2 = ddiv(Dt * t0 * (uw0 - uwm), dxe);
        pv2 += ddiv(Dt * t0 * (vw0 - vwm), dyn);
        d = d + pv2;
        d -= sig2 * t0;
    } else {
        d -= sig2 * B0[IX(g, i, j)];
    }
    out[IX(g, i, j)] = d;
}

// ------------------------------------------------ eta: hyperviscosity, AB, geofactor, update
__global__ __launch_bounds__(BX *BY) void k_eta_update(Grid g, Ptrs p, Flags *fl, double *dn, const double *m1,
                                                         const double *m2, const double *le, const double *lle,
                                                         const double *llle, int hv, double c2, double c4, double c6,
                                                         int k, double Dt, double depth) {
    CELL();
    const int q = IX(g, i, j);
    double d = dn[q];
    if (hv) {  // calc.c:346-350
        d += c2 * le[q];
        d += c4 * (lle ? lle[q] : 0.0);
        d += c6 * (llle ? llle[q] : 0.0);
        dn[q] = d;
    }
    double nab = ab(k, d, m1[q], m2[q]);
    double gf = TB(g, SW2_T_GF, j);  // Dt*sin(latc)/rZ(latc), calc.c:420
    nab -= (gf * ((p.eta[q]) + (depth)) * 0.5 * (VW(i, j) + VW(i, j - 1)));
    p.eta[q] = p.eta[q] + nab;
}

__global__ __launch_bounds__(BX *BY) void k_tracer_update(Grid g, Flags *fl, double *tr, const double *dn,
                                                            const double *m1, const double *m2, int k, int gated) {
    if (gated && !fl->nonfinite) return;
    CELL();
    const int q = IX(g, i, j);
    tr[q] = tr[q] + ab(k, dn[q], m1[q], m2[q]);
}

// ------------------------------------------------ momentum tendencies (BODYU, BODYV + pressure + drag + hypervisc + AB)
// advection of u at (i, j): BODYU, calc.c:168-200 (solve_adv_u, calc.c:424-460)
template <int SCH>
__device__ __forceinline__ double adv_u(const Grid &g, const Ptrs &p, int i, int j, double Dt) {
    double dxc = TB(g, SW2_T_DXC, j), dyv = TB(g, SW2_T_DYV, j), dxe = TB(g, SW2_T_DXE, j);
    double a0 = UW(i - 2, j), a1 = UW(i - 1, j), a2 = UW(i, j), a3 = UW(i + 1, j), a4 = UW(i + 2, j);
    double c0 = UW(i, j - 2), c1 = UW(i, j - 1), c3 = UW(i, j + 1), c4_ = UW(i, j + 2);
    double vs0 = VW(i, j - 1), vs1 = VW(i + 1, j - 1), vn0 = VW(i, j), vn1 = VW(i + 1, j);
    double CwP = ddiv(0.5 * (0.5 * (a1 + fabs(a1)) + 0.5 * (a2 + fabs(a2))) * Dt, dxc);
    double CwM = ddiv(0.5 * (0.5 * (a1 - fabs(a1)) + 0.5 * (a2 - fabs(a2))) * Dt, dxc);
    double CeP = ddiv(0.5 * (0.5 * (a2 + fabs(a2)) + 0.5 * (a3 + fabs(a3))) * Dt, dxc);
    double CeM = ddiv(0.5 * (0.5 * (a2 - fabs(a2)) + 0.5 * (a3 - fabs(a3))) * Dt, dxc);
    double CsP = ddiv(0.5 * (0.5 * (vs0 + fabs(vs0)) + 0.5 * (vs1 + fabs(vs1))) * Dt, dyv);
    double CsM = ddiv(0.5 * (0.5 * (vs0 - fabs(vs0)) + 0.5 * (vs1 - fabs(vs1))) * Dt, dyv);
    double CnP = ddiv(0.5 * (0.5 * (vn0 + fabs(vn0)) + 0.5 * (vn1 + fabs(vn1))) * Dt, dyv);
    double CnM = ddiv(0.5 * (0.5 * (vn0 - fabs(vn0)) + 0.5 * (vn1 - fabs(vn1))) * Dt, dyv);
    double rxPprev = r2rP(a0, a1, a2), rxMprev = r2rM(a1, a2, a3), rxP = r2rP(a1, a2, a3), rxM = r2rM(a2, a3, a4);
    double ryPprev = r2rP(c0, c1, a2), ryMprev = r2rM(c1, a2, c3), ryP = r2rP(c1, a2, c3), ryM = r2rM(a2, c3, c4_);
    double BwP = b2bp<SCH>(a1, a2, rxPprev, CwP), BwM = b2bm<SCH>(a1, a2, rxMprev, CwM);
    double BeP = b2bp<SCH>(a2, a3, rxP, CeP), BeM = b2bm<SCH>(a2, a3, rxM, CeM);
    double BsP = b2bp<SCH>(c1, a2, ryPprev, CsP), BsM = b2bm<SCH>(c1, a2, ryMprev, CsM);
    double BnP = b2bp<SCH>(a2, c3, ryP, CnP), BnM = b2bm<SCH>(a2, c3, ryM, CnM);
    double p1 = CwP * BwP + CwM * BwM - CeP * BeP - CeM * BeM + CsP * BsP + CsM * BsM - CnP * BnP - CnM * BnM;
    double p2 = Dt * (ddiv(a2 * (a3 - a1), (dxe + dxe)) + ddiv(a2 * ((vn0 + vn1) / 2. - (vs0 + vs1) / 2.), dyv));
    return p1 + p2;
}
// pressure gradient of u (solve_pres_u, calc.c:500-519)
__device__ __forceinline__ double pres_u(const Grid &g, const Ptrs &p, int i, int j, double Dt) {
    const int q = IX(g, i, j);
    const double dxc = TB(g, SW2_T_DXC, j);
    return ddiv(-TB(g, SW2_T_GE, j) * Dt * (p.eta[IX(g, i + 1, j)] - p.eta[q]), dxc);  // calc.c:517
}
// advection of v at (i, j): BODYV, calc.c:205-237 (solve_adv_v, calc.c:462-498)
template <int SCH>
__device__ __forceinline__ double adv_v(const Grid &g, const Ptrs &p, int i, int j, double Dt) {
    double dxv = TB(g, SW2_T_DXV, j), dyc = TB(g, SW2_T_DYC, j);
    double dyn2 = TB(g, SW2_T_DYN, j) + TB(g, SW2_T_DYN, j + 1);
    double uw_m1j1 = UW(i - 1, j + 1), uw_m1j = UW(i - 1, j), uw_j1 = UW(i, j + 1), uw_j = UW(i, j);
    double a0 = VW(i - 2, j), a1 = VW(i - 1, j), a2 = VW(i, j), a3 = VW(i + 1, j), a4 = VW(i + 2, j);
    double c0 = VW(i, j - 2), c1 = VW(i, j - 1), c3 = VW(i, j + 1), c4_ = VW(i, j + 2);
    double CwP = ddiv(0.5 * (0.5 * (uw_m1j1 + fabs(uw_m1j1)) + 0.5 * (uw_m1j + fabs(uw_m1j))) * Dt, dxv);
    double CwM = ddiv(0.5 * (0.5 * (uw_m1j1 - fabs(uw_m1j1)) + 0.5 * (uw_m1j - fabs(uw_m1j))) * Dt, dxv);
    double CeP = ddiv(0.5 * (0.5 * (uw_j1 + fabs(uw_j1)) + 0.5 * (uw_j + fabs(uw_j))) * Dt, dxv);
    double CeM = ddiv(0.5 * (0.5 * (uw_j1 - fabs(uw_j1)) + 0.5 * (uw_j - fabs(uw_j))) * Dt, dxv);
    double CsP = ddiv(0.5 * (0.5 * (c1 + fabs(c1)) + 0.5 * (a2 + fabs(a2))) * Dt, dyc);
    double CsM = ddiv(0.5 * (0.5 * (c1 - fabs(c1)) + 0.5 * (a2 - fabs(a2))) * Dt, dyc);
    double CnP = ddiv(0.5 * (0.5 * (a2 + fabs(a2)) + 0.5 * (c3 + fabs(c3))) * Dt, dyc);
    double CnM = ddiv(0.5 * (0.5 * (a2 - fabs(a2)) + 0.5 * (c3 - fabs(c3))) * Dt, dyc);
    double rxPprev = r2rP(a0, a1, a2), rxMprev = r2rM(a1, a2, a3), rxP = r2rP(a1, a2, a3), rxM = r2rM(a2, a3, a4);
    double ryPprev = r2rP(c0, c1, a2), ryMprev = r2rM(c1, a2, c3), ryP = r2rP(c1, a2, c3), ryM = r2rM(a2, c3, c4_);
    double BwP = b2bp<SCH>(a1, a2, rxPprev, CwP), BwM = b2bm<SCH>(a1, a2, rxMprev, CwM);
    double BeP = b2bp<SCH>(a2, a3, rxP, CeP), BeM = b2bm<SCH>(a2, a3, rxM, CeM);
    double BsP = b2bp<SCH>(c1, a2, ryPprev, CsP), BsM = b2bm<SCH>(c1, a2, ryMprev, CsM);
    double BnP = b2bp<SCH>(a2, c3, ryP, CnP), BnM = b2bm<SCH>(a2, c3, ryM, CnM);
    double p1 = CwP * BwP + CwM * BwM - CeP * BeP - CeM * BeM + CsP * BsP + CsM * BsM - CnP * BnP - CnM * BnM;
    double p2 = Dt * (ddiv(a2 * ((uw_j + uw_j1) / 2. - (uw_m1j + uw_m1j1) / 2.), dxv) + ddiv(a2 * (c3 - c1), dyn2));
    return p1 + p2;
}
// pressure gradient of v (solve_pres_v, calc.c:521-540)
__device__ __forceinline__ double pres_v(const Grid &g, const Ptrs &p, int i, int j, double Dt) {
    const int q = IX(g, i, j);
    return ddiv(-TB(g, SW2_T_GN, j) * Dt * (p.eta[IX(g, i, j + 1)] - p.eta[q]), TB(g, SW2_T_DYC, j + 1));  // calc.c:538
}

template <int SCH>
__global__ __launch_bounds__(BX *BY) void k_uv(Grid g, Ptrs p, Flags *fl, double Dt, double sigma, int hv, double c2,
                                                 double c4, double c6, const double *Lu, const double *Lv,
                                                 const double *LLu, const double *LLv, const double *LLLu,
                                                 const double *LLLv, double *dun, const double *um1, const double *um2,
                                                 double *dvn, const double *vm1, const double *vm2, double *duab,
                                                 double *dvab, int k) {
    CELL();
    const int q = IX(g, i, j);
    double du, dv;
    {
        double adv = adv_u<SCH>(g, p, i, j, Dt);
        double pres = pres_u(g, p, i, j, Dt);
        du = adv + pres;  // core.c:224
    }
    {
        double adv = adv_v<SCH>(g, p, i, j, Dt);
        double pres = pres_v(g, p, i, j, Dt);
        dv = adv + pres;
    }
    du -= p.u[q] * sigma * Dt;  // calc.c:396-397
    dv -= p.v[q] * sigma * Dt;
    if (hv) {  // core.c:258-266
        du += c2 * Lu[q];
        du += c4 * (LLu ? LLu[q] : 0.0);
        du += c6 * (LLLu ? LLLu[q] : 0.0);
        dv += c2 * Lv[q];
        dv += c4 * (LLv ? LLv[q] : 0.0);
        dv += c6 * (LLLv ? LLLv[q] : 0.0);
    }
    dun[q] = du;
    dvn[q] = dv;
    duab[q] = ab(k, du, um1[q], um2[q]);
    dvab[q] = ab(k, dv, vm1[q], vm2[q]);
}

// ------------------------------------------------ Coriolis (calc.c:542-598), in place
__global__ __launch_bounds__(BX *BY) void k_cor_u(Grid g, Ptrs p, Flags *fl, const double *duab, double Dt) {
    CELL();
    const int q = IX(g, i, j);
    double semiU = 0.25 * (p.v[q] + p.v[IX(g, i + 1, j)] + p.v[IX(g, i, j - 1)] + p.v[IX(g, i + 1, j - 1)]);
    double U = p.u[q];
    double Alpha = Dt * TB(g, SW2_T_FE, j) + ddiv((Dt * U * TB(g, SW2_T_SINC, j)), TB(g, SW2_T_RZC, j));
    double Beta = 0.25 * Alpha * Alpha;
    p.u[q] = ddiv((U + duab[q] - Beta * U + Alpha * semiU), (1.0 + Beta));
}

__global__ __launch_bounds__(BX *BY) void k_cor_v(Grid g, Ptrs p, Flags *fl, const double *dvab, double Dt) {
    CELL();
    const int q = IX(g, i, j);
    // U interior already holds the new values, its halo the previous ones
    double semiV = 0.25 * (p.u[IX(g, i - 1, j + 1)] + p.u[IX(g, i, j + 1)] + p.u[q] + p.u[IX(g, i - 1, j)]);
    double V = p.v[q];
    double Alpha = Dt * TB(g, SW2_T_FN, j) + ddiv((Dt * semiV * TB(g, SW2_T_SINN, j)), TB(g, SW2_T_RZN, j));
    double Beta = 0.25 * Alpha * Alpha;
    p.v[q] = ddiv((V + dvab[q] - Beta * V - Alpha * semiV), (1.0 + Beta));
}

// ------------------------------------------------ Laplacians (laplacian.c:51-151)
__global__ __launch_bounds__(BX *BY) void k_lapv_zd(Grid g, Flags *fl, const double *uin, const double *vin, double *ze,
                                                      double *di) {
    CELL();
    const int q = IX(g, i, j);
    double a = TB(g, SW2_T_IDXV, j), b = -(TB(g, SW2_T_IDYC, j + 1));
    ze[q] = a * ((vin[IX(g, i + 1, j)] - vin[q]) +
                 b * (TB(g, SW2_T_DXC, j + 1) * uin[IX(g, i, j + 1)] - TB(g, SW2_T_DXC, j) * uin[q]));
    a = TB(g, SW2_T_IDXE, j);
    b = TB(g, SW2_T_IDYN, j);
    di[q] = a * ((uin[q] - uin[IX(g, i - 1, j)]) + b * (TB(g, SW2_T_DXV, j) * vin[q] - TB(g, SW2_T_DXV, j - 1) * vin[IX(g, i, j - 1)]));
}

__global__ void k_lapv_fix1(Grid g, Flags *fl, const double *uin, double *ze, double *di) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step || i > g.nx) return;
    di[IX(g, i, g.ny + 1)] = ddiv((uin[IX(g, i, g.ny)] - uin[IX(g, i - 1, g.ny)]), TB(g, SW2_T_DXE, g.ny));
    ze[IX(g, i, 0)] = 0;
}

__global__ __launch_bounds__(BX *BY) void k_lapv_out(Grid g, Flags *fl, const double *ze, const double *di, double *uo,
                                                       double *vo) {
    CELL();
    const int q = IX(g, i, j);
    double a = -(TB(g, SW2_T_IDYN, j)), b = TB(g, SW2_T_IDXC, j);
    uo[q] = a * (ze[q] - ze[IX(g, i, j - 1)]) + b * (di[IX(g, i + 1, j)] - di[q]);
    a = TB(g, SW2_T_IDXV, j);
    b = TB(g, SW2_T_IDYC, j + 1);
    vo[q] = a * (ze[q] - ze[IX(g, i - 1, j)]) + b * (di[IX(g, i, j + 1)] - di[q]);
}

__global__ void k_lapv_fix2(Grid g, Flags *fl, double *uo, double *vo) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step || i > g.nx) return;
    uo[IX(g, i, g.ny + 1)] = uo[IX(g, i, g.ny)];
    vo[IX(g, i, g.ny + 1)] = 0;
    uo[IX(g, i, 0)] = uo[IX(g, i, 1)];
    vo[IX(g, i, 0)] = 0;
}

__global__ __launch_bounds__(BX *BY) void k_lape_px(Grid g, Flags *fl, const double *ein, double *psi_, double *xi) {
    CELL();
    const int q = IX(g, i, j);
    double fp = TB(g, SW2_T_FP, j), fx = TB(g, SW2_T_FX, j);
    psi_[q] = ddiv(fp * (ein[IX(g, i + 1, j)] - ein[q]), (TB(g, SW2_T_DXC, j)));
    xi[q] = ddiv(fx * (ein[IX(g, i, j + 1)] - ein[q]), (TB(g, SW2_T_DYC, j)));
}

__global__ __launch_bounds__(BX *BY) void k_lape_out(Grid g, Flags *fl, const double *psi_, const double *xi, double *eo) {
    CELL();
    const int q = IX(g, i, j);
    double kk = TB(g, SW2_T_KK, j);
    eo[q] = kk * (ddiv((psi_[q] - psi_[IX(g, i - 1, j)]), (TB(g, SW2_T_DXE, j))) + ddiv((xi[q] - xi[IX(g, i, j - 1)]), (TB(g, SW2_T_DYN, j))));
}

__global__ void k_lape_fix(Grid g, Flags *fl, double *eo) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step || i > g.nx) return;
    eo[IX(g, i, g.ny + 1)] = eo[IX(g, i, g.ny)];
    eo[IX(g, i, g.ny + 2)] = eo[IX(g, i, g.ny + 1)];
    eo[IX(g, i, 0)] = eo[IX(g, i, 1)];
}


// single-map halo_update: every halo cell takes the value of its periodic
// image in x and y (y first for interior columns, then x for all rows)
__global__ void k_halo(Grid g, Flags *fl, double *f, int gated) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step || (gated && !fl->nonfinite)) return;
    const int W = g.nx + 4;
    int i, j;
    if (t < 4 * W) {
        int r = t / W;
        i = (t % W) - 1;
        j = r < 2 ? r - 1 : g.ny + (r - 1);
    } else {
        int t2 = t - 4 * W;
        if (t2 >= 4 * g.ny) return;
        j = 1 + t2 / 4;
        int c = t2 % 4;
        i = c < 2 ? c - 1 : g.nx + (c - 1);
    }
    f[IX(g, i, j)] = f[IX(g, wrap(i, g.nx), wrap(j, g.ny))];
}

// do_channel for u and v (windx, windy are fixed after the first call)
__global__ void k_channel(Grid g, Flags *fl, double *u, double *v) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step || i > g.nx) return;
    const int ny = g.ny;
    v[IX(g, i, 0)] = 0;
    v[IX(g, i, -1)] = 0;
    u[IX(g, i, 0)] = u[IX(g, i, 1)];
    u[IX(g, i, -1)] = u[IX(g, i, 1)];
    v[IX(g, i, ny)] = 0;
    v[IX(g, i, ny + 1)] = 0;
    v[IX(g, i, ny + 2)] = 0;
    u[IX(g, i, ny + 1)] = u[IX(g, i, ny)];
    u[IX(g, i, ny + 2)] = u[IX(g, i, ny)];
}


// the checks of one step all run, even after one of them failed: the step
// completes, as SW2's does before check_stats_scaf aborts
__global__ __launch_bounds__(BX *BY) void k_check(Grid g, Flags *fl, const double *f, int n, int kind, int gated) {
    if (gated && !fl->nonfinite) return;
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.y * BY + threadIdx.y;
    if (fl->fail_step && fl->fail_step != n) return;
    if (i > g.nx || j > g.ny) return;
    double x = f[IX(g, i, j)];
    if ((kind == 4 || kind == 5) && !isfinite(x)) fl->nonfinite = 1;
    bool bad = (x < -1e10) || (x > 1e10) || (i == 1 && j == 1 && x != x);
    if (bad) flag_bad(fl, n, kind);
}

// ------------------------------------------------ in-loop perturbations
__global__ void k_add(Grid g, Flags *fl, double *f, const double *inc) {
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.y * BY + threadIdx.y;
    if (fl->fail_step || i > g.nx || j > g.ny) return;
    const int q = IX(g, i, j);
    f[q] += inc[q];
}

// one storm injection from its sparse list (the cells inside the radius,
// perturbs.c:348-395): the tracer value and the perturbation mark there, and
// the eta increment when the injection changes eta
// SW2 injects before the step's checks (core.c:280-312): a check failing in
// this step (the fused step checks u and v before the injection) leaves it
// to run; an earlier failure or this step's CFL guard does not
__global__ void k_inject(Flags *fl, int step, int n, const int *idx, const double *inc, const double *trv, int has_inc,
                         double *eta, double *tr, double *pert) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ((fl->fail_step && (fl->fail_step != step || fl->fail_kind == 1)) || k >= n) return;
    const int q = idx[k];
    tr[q] = trv[k];
    pert[q] = 1.0;
    if (has_inc) eta[q] += inc[k];
}

// clear the perturbation marks of the previous injections at step `step`;
// like the injection itself, not once SW2 has stopped (an earlier failure,
// or this step's CFL guard), so the marks of the failing step remain
__global__ void k_zero_pert(Flags *fl, int step, int S, double *f) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step && (fl->fail_step != step || fl->fail_kind == 1)) return;
    if (q < S) f[q] = 0;
}

}  // namespace

// Tiles of 32 x 16 cells: every single-run kernel and the default batch
// kernels.  A 32 x 16 tile stages 37 x 21 = 777 values per array for its 512
// cells, where a 64 x 8 tile stages 69 x 13 = 897: fewer halo rows per cell
// for the staging and the Laplacian passes (measured on the 4096 x 2048
// workload, MI355X: FP64 step 0.833 -> 0.786 ms, mixed step 0.515 -> 0.430 ms;
// the 64-member ensemble 2.45 -> 2.00 ms per step).  Each tile shape's
// helpers and kernels live in a namespace of their own, so that
// argument-dependent lookup (Grid and the argument records live in the
// anonymous namespace) never sees two shapes' functions at once.  The batch
// kernels, of every tile shape, are compiled in sw2_batch.hip, once per
// shape, and launched through sw2_batch_phases_<shape>.
#define SW2_TILE_X 32
#define SW2_TILE_Y 16
namespace tile_32x16 {
#include "sw2_tile.inc"
}  // namespace tile_32x16
using namespace tile_32x16;
#include "sw2_fused.inc"
#include "sw2_mixed_common.inc"
namespace tile_32x16 {
#include "sw2_mixed.inc"
}  // namespace tile_32x16
#undef SW2_TILE_X
#undef SW2_TILE_Y

namespace {
// storm injections on the mixed form (perturbs.c:348-395): the marks as
// k_inject sets them, the eta increment on the height anomaly eta' (eta =
// hbar(j) + eta') and, for storms that inject tracer, the tracer value
// (tr_on), each written with its periodic images as the kernels write them
__device__ __forceinline__ void inject_cell(const Grid &g, int q, double inc, double trv, int has_inc, MField eta, MField tr,
                                            int tr_on, double *pert) {
    pert[q] = 1.0;
    const int c = q % g.sx - 1, j = q / g.sx - 1;
    if (tr_on) mput_periodic<float, true>(g, tr.hi, c, j, mset_lo(tr, q, trv));
    if (!has_inc) return;
    mput_periodic<float, true>(g, eta.hi, c, j, mset_lo(eta, q, mget(eta, q) + inc));
}

__global__ void k_inject_mixed(Grid g, Flags *fl, int step, int n, const int *idx, const double *inc, const double *trv,
                               int has_inc, MField eta, MField tr, int tr_on, double *pert) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ((fl->fail_step && (fl->fail_step != step || fl->fail_kind == 1)) || k >= n) return;
    inject_cell(g, idx[k], inc[k], trv[k], has_inc, eta, tr, tr_on, pert);
}

// check_stats_scaf on a mixed form field (hbar(j) + the anomaly, for eta):
// the checks the momentum kernel leaves for after the storms
__global__ __launch_bounds__(BX *BY) void k_check_mixed(Grid g, Flags *fl, MField f, const double *hbar, int n, int kind) {
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.y * BY + threadIdx.y;
    if (fl->fail_step && fl->fail_step != n) return;
    if (i > g.nx || j > g.ny) return;
    const double x = (hbar ? hbar[j + 1] : 0.0) + mget(f, IX(g, i, j));
    if ((x < -1e10) || (x > 1e10) || (i == 1 && j == 1 && x != x)) flag_bad(fl, n, kind);
}

// a vortex's increments on a mixed form field (perturbs.c:397-460; eta' for
// eta, whose mean does not change): value and residual of the interior
// cells, their halo images by k_halo_mixed after every vortex of the step
__global__ __launch_bounds__(BX *BY) void k_add_mixed(Grid g, Flags *fl, MField f, const double *inc) {
    const int i = 1 + blockIdx.x * BX + threadIdx.x;
    const int j = 1 + blockIdx.y * BY + threadIdx.y;
    if (fl->fail_step || i > g.nx || j > g.ny) return;
    const int q = IX(g, i, j);
    f.hi[q] = mset_lo(f, q, mget(f, q) + inc[q]);
}

// a mixed form field in FP64 (value + residual on the interior, the value
// on the halo), the per-phase Laplacian kernels' input in mode 3
__global__ void k_mixed_f64(Grid g, MField f, double *out, int S) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= S) return;
    const int i = q % g.sx - 1, j = q / g.sx - 1;
    const bool interior = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny;
    out[q] = interior ? mget(f, q) : (double)f.hi[q];
}

// halo_update (sppde_hupdate.c:274-363) on the values of a mixed form field,
// as k_halo does on an FP64 field: every halo cell the periodic image
__global__ void k_halo_mixed(Grid g, Flags *fl, float *f) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step) return;
    const int W = g.nx + 4;
    int i, j;
    if (t < 4 * W) {
        const int r = t / W;
        i = (t % W) - 1;
        j = r < 2 ? r - 1 : g.ny + (r - 1);
    } else {
        const int t2 = t - 4 * W;
        if (t2 >= 4 * g.ny) return;
        j = 1 + t2 / 4;
        const int c = t2 % 4;
        i = c < 2 ? c - 1 : g.nx + (c - 1);
    }
    f[IX(g, i, j)] = f[IX(g, wrap(i, g.nx), wrap(j, g.ny))];
}

// do_channel (calc.c:600-649) on the mixed form's u and v, as k_channel:
// interior columns only; the wall row ny of v is an interior row, so its
// residual is cleared with its value
__global__ void k_channel_mixed(Grid g, Flags *fl, MField u, MField v) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (fl->fail_step || i > g.nx) return;
    const int ny = g.ny;
    v.hi[IX(g, i, 0)] = 0.f;
    v.hi[IX(g, i, -1)] = 0.f;
    u.hi[IX(g, i, 0)] = u.hi[IX(g, i, 1)];
    u.hi[IX(g, i, -1)] = u.hi[IX(g, i, 1)];
    v.hi[IX(g, i, ny)] = 0.f;
    v.lo[IX(g, i, ny)] = 0.f;
    v.hi[IX(g, i, ny + 1)] = 0.f;
    v.hi[IX(g, i, ny + 2)] = 0.f;
    u.hi[IX(g, i, ny + 1)] = u.hi[IX(g, i, ny)];
    u.hi[IX(g, i, ny + 2)] = u.hi[IX(g, i, ny)];
}

// Storms of batch members inside a replayed graph.  The host evaluates a
// window of steps ahead (the cells of every storm at each step, as a single
// session does step by step) and writes the window table; the graph's three
// kernels per step read their row of it.  They do what k_zero_pert,
// k_inject_mixed and k_check_mixed do for a single mixed session, with the
// same gates and the same operations per cell, in the same order, so that a
// member's result is the one it has when run alone.
constexpr int SMAXK = 8;  // storm lists per member and step (sw2_gaussians_sparse's maxk)
struct SEntry {           // one storm member at one step of the window
    Grid g;
    Flags *fl;
    int n;        // the step; 0: the member takes no step here
    int zero, S;  // clear the marks of the previous injections (all S cells) first
    double *pert;
    MField eta, tr;  // the step's new height anomaly, and the tracer
    int tr_on;
    const double *hbar;
    int nk;  // storm lists of this step
    int cnt[SMAXK], has_inc[SMAXK];
    const int *idx[SMAXK];
    const double *inc[SMAXK], *trv[SMAXK];
};

// the marks (grid: blocks x entries)
__global__ __launch_bounds__(256) void kbs_zero(const SEntry *__restrict__ w) {
    const SEntry &E = w[blockIdx.y];
    if (!E.n || !E.zero) return;
    const Flags *fl = E.fl;
    if (fl->fail_step && (fl->fail_step != E.n || fl->fail_kind == 1)) return;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < E.S; q += gridDim.x * 256) E.pert[q] = 0;
}

// the injections: one workgroup per entry, its lists one after another as
// the single session's launches run them (a cell in two storms takes the
// second after the first)
__global__ __launch_bounds__(1024) void kbs_inject(const SEntry *__restrict__ w) {
    const SEntry &E = w[blockIdx.y];
    if (!E.n) return;
    const Flags *fl = E.fl;
    if (fl->fail_step && (fl->fail_step != E.n || fl->fail_kind == 1)) return;
    for (int k = 0; k < E.nk; k++) {
        for (int c = threadIdx.x; c < E.cnt[k]; c += 1024)
            inject_cell(E.g, E.idx[k][c], E.inc[k][c], E.trv[k][c], E.has_inc[k], E.eta, E.tr, E.tr_on, E.pert);
        __syncthreads();
    }
}

// the tracer and height checks after the injections
__global__ __launch_bounds__(256) void kbs_check(const SEntry *__restrict__ w) {
    const SEntry &E = w[blockIdx.y];
    if (!E.n) return;
    Flags *fl = E.fl;
    if (fl->fail_step && fl->fail_step != E.n) return;
    const Grid g = E.g;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < g.nx * g.ny; q += gridDim.x * 256) {
        const int i = 1 + q % g.nx, j = 1 + q / g.nx;
        if (E.tr_on) {
            const double x = 0.0 + mget(E.tr, IX(g, i, j));
            if ((x < -1e10) || (x > 1e10) || (i == 1 && j == 1 && x != x)) flag_bad(fl, E.n, 2);
        }
        const double x = E.hbar[j + 1] + mget(E.eta, IX(g, i, j));
        if ((x < -1e10) || (x > 1e10) || (i == 1 && j == 1 && x != x)) flag_bad(fl, E.n, 3);
    }
}

// device reductions of SW2's energy diagnostics (energy.c:25-55) and of the
// mass sum H*Dxc*Dyc over the interior, in FP64, for either form: per-block
// partial sums, then one block adding them in a fixed order
// (the mixed form: T float with the residual arrays lo_*, which hold the
// interior cells; the reduction reads u at i-1 = 0 and v at j-1 = 0, whose
// values are the images of (nx, j) and the wall's zero)
template <typename T>
__global__ __launch_bounds__(256) void k_energy(Grid g, const T *eta, const T *u, const T *v, const double *hbar,
                                                  double depth, double *part, const float *lo_eta = nullptr,
                                                  const float *lo_u = nullptr, const float *lo_v = nullptr) {
    double ke = 0, ape = 0, mass = 0;
    auto val = [&](const T *f, const float *lo, int i, int j) -> double {
        const int q = IX(g, i, j);
        if (!lo) return (double)f[q];
        const int iw = i < 1 ? i + g.nx : i;  // the interior cell whose residual the image shares
        return (j < 1) ? (double)f[q] : (double)f[q] + (double)lo[IX(g, iw, j)];
    };
    for (int q = blockIdx.x * 256 + threadIdx.x; q < g.nx * g.ny; q += gridDim.x * 256) {
        const int i = 1 + q % g.nx, j = 1 + q / g.nx;
        const double wx = TB(g, T_WROW, j);
        const double H = val(eta, lo_eta, i, j) + (hbar ? hbar[j + 1] : 0.0) + depth;
        const double uc = 0.5 * ((val(u, lo_u, i, j) + wx) + (val(u, lo_u, i - 1, j) + wx));
        const double vc = 0.5 * (val(v, lo_v, i, j) + val(v, lo_v, i, j - 1));
        const double surf = TB(g, SW2_T_DXC, j) * TB(g, SW2_T_DYC, j), gc = TB(g, SW2_T_GC, j);
        ke += 0.5 * H * (uc * uc + vc * vc) * surf - 0.5 * (wx * wx) * surf;
        ape += 0.5 * gc * (H * H) * surf - 0.5 * gc * (depth * depth) * surf;
        mass += H * surf;
    }
    __shared__ double r[3][256];
    r[0][threadIdx.x] = ke;
    r[1][threadIdx.x] = ape;
    r[2][threadIdx.x] = mass;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) r[k][threadIdx.x] += r[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[3 * blockIdx.x + k] = r[k][0];
}

__global__ void k_energy_fin(int nb, const double *part, double *out) {
    if (threadIdx.x != 0) return;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int b = 0; b < nb; b++) {
        s0 += part[3 * b];
        s1 += part[3 * b + 1];
        s2 += part[3 * b + 2];
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
}
}  // namespace

// ------------------------------------------------ energy without allocation
// sw2_session_energy_noalloc keeps the partial sums of the energy reduction
// (ENERGY_PARTS blocks x KE, APE, mass) in this file-scope array instead of
// the session's scratch: every device the library runs on holds one copy
// until the process ends.  In host code `scratch` names a host-side stand-in,
// never the device copy: the device address comes from hipGetSymbolAddress,
// taken once per device with that device current and kept in scratch_dev[].
constexpr int ENERGY_PARTS = 1024;  // blocks of k_energy: any grid reduces to this many partials
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[3 * ENERGY_PARTS];
static double *scratch_dev[MAX_DEVICES];

// the device address of `scratch` on the current device (taken on the first
// call per device; sw2_session_create makes that call, so a caller that
// captures graphs has it before capture begins)
static int scratch_address(double **out) {
    int dev = -1;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        sw2_set_error("device %d: the scratch table holds devices 0..%d", dev, MAX_DEVICES - 1);
        return -1;
    }
    double *p = __atomic_load_n(&scratch_dev[dev], __ATOMIC_ACQUIRE);
    if (!p) {
        void *a = nullptr;
        HIPCHECK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        p = (double *)a;
        __atomic_store_n(&scratch_dev[dev], p, __ATOMIC_RELEASE);
    }
    *out = p;
    return 0;
}

// ====================================================================== session
struct sw2_session {
    sw *SW;
    Grid g;
    int S, form, sch, hv, tracer_active, nstart, n, perturbs_dirty;
    int fused;      // 0: one kernel per phase; 1: the two fused kernels per step
    int hvmode;     // fused path: 0 none, 2 del-2 and del-4 in LDS, 3 Laplacians from the per-phase kernels
    int cfl_ready;  // CFL maxima of the next step already reduced
    double *eta2, *u2, *v2;  // ping-pong partners of eta, u, v on the fused path
    // mixed form: FP32 anomaly state, its ping-pong partners and histories;
    // FP64 mean-state tables (hbar, H0, mean pressure gradient)
    MField m_eta, m_u, m_v, m_eta2, m_u2, m_v2;
    float *m_De[3], *m_Du[3], *m_Dv[3];
    MField m_tr;      // the tracer (tracer_active), updated in place
    float *m_Dtr[3];  // and its tendencies
    int mhv;          // the mixed kernels' hyperviscosity mode: 0, 2 (del-2/del-4 in LDS), 3 (del-6)
    double *m_u64, *m_v64;  // mode 3: u, v in FP64, the per-phase Laplacian kernels' input
    double *d_mt, *h_hbar;
    double *d_part;          // energy reduction scratch
    unsigned long long *d_cflp;  // partial CFL maxima when stepped in a batch
    hipEvent_t *kev;         // per-kernel timing (sw2_session_kernel_times): 3 events a step
    int kev_i;
    double t, Dt, depth, sigma, tracer_dis, courant, rest_x, rest_y;
    double c2, c4, c6;
    double *d_tab;
    double *eta, *u, *v, *tr, *pert, *wx, *wy;
    double *Deta[3], *Du[3], *Dv[3], *Dtr[3];
    double *duab, *dvab;
    double *Lu, *Lv, *LLu, *LLv, *LLLu, *LLLv, *ze, *di, *le, *lle, *llle, *psi, *xi;
    double *pinc, *pdu, *pdv;
    Flags *fl;
    int fl_borrowed;  // fl lies in a batch's array of guard records (the batch frees it)
    hipStream_t st;
    // host scratch for perturbations: storm injections as sparse lists
    // (inj, uploaded to d_idx/d_inc/d_trv), vortices as whole fields
    sw2_inj inj[8];
    int *d_idx[8], d_cap[8];
    double *d_inc[8], *d_trv[8];
    double *h_dv[3][8];
    // pinned staging of the lists, two sets used alternately: a set is
    // refilled once the event after its copies has passed, so the host
    // builds the next step's lists while the device still runs this one
    char *h_stage[2];
    size_t stage_cap[2];
    hipEvent_t stage_ev[2];
    int stage_i;
};

static int dalloc(double **p, size_t n) {
    HIPCHECK(hipMalloc((void **)p, n * sizeof(double)));
    HIPCHECK(hipMemset(*p, 0, n * sizeof(double)));
    return 0;
}

static dim3 cells_grid(const Grid &g) { return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY); }
static dim3 cells_block() { return dim3(BX, BY); }

static void rot(double *h[3]) {
    double *m2 = h[2];
    h[2] = h[1];
    h[1] = h[0];
    h[0] = m2;
}

static int falloc(float **p, size_t n) {
    HIPCHECK(hipMalloc((void **)p, n * sizeof(float)));
    HIPCHECK(hipMemset(*p, 0, n * sizeof(float)));
    return 0;
}

// the mixed form's mean state and anomaly fields from the world
static int mixed_setup(sw2_session *s, sw *SW) {
    const int R = s->g.R, nx = SW->nx;
    const size_t S = (size_t)s->S;
    map *M = SW->M;
    double *mt = (double *)calloc((size_t)MT_N * R, sizeof(double));
    s->h_hbar = (double *)calloc((size_t)R, sizeof(double));
    for (int j = -1; j <= SW->ny + 2; j++) { /* zonal mean of every array row */
        double m = 0;
        for (int i = 1; i <= nx; i++) m += ac(SW->eta, i, j, M);
        s->h_hbar[j + 1] = m / nx;
    }
    for (int r = 0; r < R; r++) {
        mt[MT_HBAR * R + r] = s->h_hbar[r];
        mt[MT_H0 * R + r] = SW->depth + s->h_hbar[r];
        if (r + 1 < R) { /* -gn*Dt*(hbar(j+1)-hbar(j))/Dyc(j+1), the mean share of Dv_pres (calc.c:538) */
            int j = r - 1;
            mt[MT_P0 * R + r] = -ac(SW->gn, 1, j, M) * SW->Dt * (s->h_hbar[r + 1] - s->h_hbar[r]) / ac(SW->Dyc, 1, j + 1, M);
        }
    }
    int rc = dalloc(&s->d_mt, (size_t)MT_N * R);
    if (!rc && hipMemcpy(s->d_mt, mt, sizeof(double) * MT_N * R, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    free(mt);
    MField *fs[] = {&s->m_eta, &s->m_u, &s->m_v, &s->m_eta2, &s->m_u2, &s->m_v2};
    for (auto f : fs) {
        rc |= falloc(&f->hi, S);
        rc |= falloc(&f->lo, S);
    }
    if (!rc && SW->num_Gaussians) { /* the storms' marks, as SW2's perturbs field */
        rc = dalloc(&s->pert, S);
        if (!rc && hipMemcpy(s->pert, SW->perturbs, sizeof(double) * S, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    }
    for (int k = 0; k < 3; k++) {
        rc |= falloc(&s->m_De[k], S);
        rc |= falloc(&s->m_Du[k], S);
        rc |= falloc(&s->m_Dv[k], S);
    }
    if (s->tracer_active) {
        rc |= falloc(&s->m_tr.hi, S);
        rc |= falloc(&s->m_tr.lo, S);
        for (int k = 0; k < 3; k++) rc |= falloc(&s->m_Dtr[k], S);
    }
    // del-6: the three Laplacian passes of u and v by the per-phase FP64
    // kernels, from an FP64 copy of u and v, before each step
    s->mhv = !s->hv ? 0 : (SW->nu6 != 0 ? 3 : 2);
    if (s->mhv == 3) {
        double **hvf[] = {&s->m_u64, &s->m_v64, &s->Lu, &s->Lv, &s->LLu, &s->LLv, &s->LLLu, &s->LLLv,
                          &s->ze, &s->di, &s->le, &s->lle, &s->llle, &s->psi, &s->xi};
        for (auto f : hvf) rc |= dalloc(f, S);
    }
    if (rc) return -1;
    // each field split into its FP32 value and residual
    float *hh = (float *)malloc(sizeof(float) * S), *hl = (float *)malloc(sizeof(float) * S);
    const double *src[4] = {SW->eta, SW->u, SW->v, SW->tracer};
    const MField dst[4] = {s->m_eta, s->m_u, s->m_v, s->m_tr};
    for (int f = 0; f < (s->tracer_active ? 4 : 3) && !rc; f++) {
        for (int j = -1; j <= SW->ny + 2; j++)
            for (int i = -1; i <= nx + 2; i++) {
                int q = LP(i, j, M);
                const double x = src[f][q] - (f == 0 ? s->h_hbar[j + 1] : 0.0);
                hh[q] = (float)x;
                hl[q] = (float)(x - (double)hh[q]);
            }
        if (hipMemcpy(dst[f].hi, hh, sizeof(float) * S, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dst[f].lo, hl, sizeof(float) * S, hipMemcpyHostToDevice) != hipSuccess)
            rc = -1;
    }
    free(hh);
    free(hl);
    if (rc) sw2_set_error("upload failed");
    return rc;
}

static void frot(float *h[3]) {
    float *m2 = h[2];
    h[2] = h[1];
    h[1] = h[0];
    h[0] = m2;
}

// fl_at: where the session's guard record lives when a batch owns it (a
// slot of the batch's array); NULL: the session allocates its own
static int session_create(sw *SW, int form, sw2_session **out, Flags *fl_at) {
    *out = NULL;
    const int force_phased = (form & SW2_PHASED) != 0;
    form &= ~SW2_PHASED;
    if (form != SW2_FORM_FP64 && form != SW2_FORM_MIXED) {
        sw2_set_error("form %d is not available", form);
        return -1;
    }
    if (SW->polar != 0) {
        sw2_set_error("only the channel geometry (polar 0) runs on the GPU");
        return -1;
    }
    if (!sw2_all_zero(SW->hB, SW->M->S)) {
        sw2_set_error("hB must be identically zero (SW2 never sets it)");
        return -1;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        sw2_set_error("no HIP device available");
        return -1;
    }
    sw2_session *s = (sw2_session *)calloc(1, sizeof(sw2_session));
    s->SW = SW;
    s->S = SW->M->S;
    s->sch = SW->sch;
    s->form = form;
    s->Dt = SW->Dt;
    s->depth = SW->depth;
    s->sigma = SW->sigma_dis;
    s->tracer_dis = SW->tracer_dis;
    s->courant = SW->Courant_max;
    s->hv = (SW->nu2 != 0 || SW->nu4 != 0 || SW->nu6 != 0);
    s->c2 = (SW->nu2) * (SW->Dt);
    s->c4 = (SW->nu4) * (SW->Dt);
    s->c6 = (SW->nu6) * (SW->Dt);
    sw2_cfl_rest(SW, &s->rest_x, &s->rest_y);
    int inject = 0;
    for (int k = 0; k < SW->num_Gaussians; k++)
        if (SW->Gaussians[k * N_COLS_GAUSS + 11] != 0) inject = 1;
    s->tracer_active = inject || !sw2_all_zero(SW->tracer, s->S);
    s->g.nx = SW->nx;
    s->g.ny = SW->ny;
    s->g.sx = SW->nx + 4;
    s->g.R = SW->ny + 4;
    size_t S = (size_t)s->S;
    int rc = 0;
    const int R = s->g.R, NTD = SW2_NT + 2;
    double *tab = (double *)calloc((size_t)NTD * R, sizeof(double));
    const int tables_ok = sw2_row_tables(SW, tab);
    // windx as the fused kernels read it: the row value everywhere except in
    // the four corner blocks, which keep the y-wrap of the first halo update
    int wind_rows_ok = 1;
    for (int j = -1; j <= SW->ny + 2; j++) {
        tab[T_WROW * R + j + 1] = ac(SW->windx, 1, j, SW->M);
        tab[T_WCOR * R + j + 1] = ac(SW->windx, 0, j, SW->M);
        for (int i = -1; i <= SW->nx + 2; i++) {
            bool corner = (i < 1 || i > SW->nx) && (j < 1 || j > SW->ny);
            double w = corner ? tab[T_WCOR * R + j + 1] : tab[T_WROW * R + j + 1];
            if (memcmp(&w, &ac(SW->windx, i, j, SW->M), sizeof w) != 0) wind_rows_ok = 0;
        }
    }
    int windy_zero = 1;
    for (int q = 0; q < s->S; q++) {
        unsigned long long b;
        memcpy(&b, &SW->windy[q], sizeof b);
        if (b) windy_zero = 0;
    }
    s->fused = !force_phased && tables_ok && wind_rows_ok && windy_zero && SW->nx >= 128 && SW->ny >= 4;
    // del-2 with del-4 forms both passes in the kernels' LDS; del-6, or del-2
    // alone, takes the Laplacians from the per-phase kernels (mode 3)
    s->hvmode = !s->hv ? 0 : ((SW->nu6 == 0 && !(SW->nu2 != 0 && SW->nu4 == 0)) ? 2 : 3);
    // the mixed kernels form del-2 and del-4 in LDS; del-2 alone runs there
    // with a zero del-4 coefficient (the FP64 step takes it from the
    // per-phase kernels only to leave SW2's sum without the del-4 term bit
    // for bit); del-6 takes its three Laplacian passes from the per-phase
    // FP64 kernels, run on an FP64 copy of u and v (mixed_setup, step_mixed)
    if (form == SW2_FORM_MIXED && !s->fused) {
        sw2_set_error("the mixed-precision form needs a case the fused step takes: a j-only wind, nx >= 128, ny >= 4");
        free(tab);
        free(s);
        return -1;
    }
    rc |= dalloc(&s->d_tab, (size_t)NTD * R);
    if (!rc && hipMemcpy(s->d_tab, tab, sizeof(double) * NTD * R, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    free(tab);
    s->g.T = s->d_tab;
    rc |= dalloc(&s->d_part, 3 * ENERGY_PARTS + 3);
    // the momentum kernels' partial CFL maxima, zero between steps
    if (!rc && hipMalloc((void **)&s->d_cflp, sizeof(unsigned long long) * 2 * NPART) != hipSuccess) rc = -1;
    if (!rc && hipMemset(s->d_cflp, 0, sizeof(unsigned long long) * 2 * NPART) != hipSuccess) rc = -1;
    {
        double *sp = nullptr;  // the allocation-free energy's scratch address on this device
        if (!rc && scratch_address(&sp)) rc = -1;
    }
    if (form == SW2_FORM_MIXED) {
        if (!rc) rc = mixed_setup(s, SW);
        if (fl_at) {
            s->fl = fl_at;
            s->fl_borrowed = 1;
        } else if (!rc && hipMalloc((void **)&s->fl, sizeof(Flags)) != hipSuccess) {
            rc = -1;
        }
        if (!rc && hipMemset(s->fl, 0, sizeof(Flags)) != hipSuccess) rc = -1;
        if (!rc && hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) rc = -1;
        if (rc) {
            if (!sw2_last_error()[0]) sw2_set_error("device allocation failed");
            sw2_session_destroy(s);
            return -1;
        }
        s->t = SW->t0;
        *out = s;
        return 0;
    }
    if (s->fused) {
        rc |= dalloc(&s->eta2, S);
        rc |= dalloc(&s->u2, S);
        rc |= dalloc(&s->v2, S);
    }
    double **fields[] = {&s->eta, &s->u, &s->v, &s->tr, &s->pert, &s->wx, &s->wy, &s->duab, &s->dvab};
    for (auto f : fields) rc |= dalloc(f, S);
    for (int k = 0; k < 3; k++) {
        rc |= dalloc(&s->Deta[k], S);
        rc |= dalloc(&s->Du[k], S);
        rc |= dalloc(&s->Dv[k], S);
        rc |= dalloc(&s->Dtr[k], S);
    }
    if (s->hv) {
        double **hvf[] = {&s->Lu, &s->Lv, &s->ze, &s->di, &s->le, &s->psi, &s->xi};
        for (auto f : hvf) rc |= dalloc(f, S);
        if (SW->nu4 != 0 || SW->nu6 != 0) {
            rc |= dalloc(&s->LLu, S);
            rc |= dalloc(&s->LLv, S);
            rc |= dalloc(&s->lle, S);
        }
        if (SW->nu6 != 0) {
            rc |= dalloc(&s->LLLu, S);
            rc |= dalloc(&s->LLLv, S);
            rc |= dalloc(&s->llle, S);
        }
    }
    if (!rc && hipMalloc((void **)&s->fl, sizeof(Flags)) != hipSuccess) rc = -1;
    if (!rc && hipMemset(s->fl, 0, sizeof(Flags)) != hipSuccess) rc = -1;
    if (!rc && hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) rc = -1;
    if (rc) {
        if (!sw2_last_error()[0]) sw2_set_error("device allocation failed");
        sw2_session_destroy(s);
        return -1;
    }
    struct {
        double *d;
        double *h;
    } up[] = {{s->eta, SW->eta}, {s->u, SW->u}, {s->v, SW->v}, {s->tr, SW->tracer}, {s->pert, SW->perturbs}, {s->wx, SW->windx}, {s->wy, SW->windy}};
    for (auto &x : up)
        if (hipMemcpy(x.d, x.h, sizeof(double) * S, hipMemcpyHostToDevice) != hipSuccess) {
            sw2_set_error("upload failed");
            sw2_session_destroy(s);
            return -1;
        }
    s->perturbs_dirty = !sw2_all_zero(SW->perturbs, s->S);
    s->t = SW->t0;
    *out = s;
    return 0;
}

extern "C" int sw2_session_create(sw *SW, int form, sw2_session **out) { return session_create(SW, form, out, NULL); }

extern "C" int sw2_session_begin(sw2_session *s, int nstart, double t) {
    s->nstart = nstart;
    s->cfl_ready = 0;
    s->n = nstart;
    s->t = t;
    HIPCHECK(hipMemsetAsync(s->fl, 0, sizeof(Flags), s->st));
    HIPCHECK(hipStreamSynchronize(s->st));
    return 0;
}

static void launch_halo(sw2_session *s, double *f, int gated = 0) {
    int cnt = 4 * (s->g.nx + 4) + 4 * s->g.ny;
    k_halo<<<(cnt + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, f, gated);
}

template <int SCH>
static void launch_flux(sw2_session *s, Ptrs p, bool tracer, const double *B, double sig2, double *out) {
    if (tracer)
        k_flux<SCH, true><<<cells_grid(s->g), cells_block(), 0, s->st>>>(s->g, p, s->fl, B, s->depth, s->Dt, sig2, out,
                                                                         !s->tracer_active);
    else k_flux<SCH, false><<<cells_grid(s->g), cells_block(), 0, s->st>>>(s->g, p, s->fl, B, s->depth, s->Dt, sig2, out, 0);
}

template <int SCH>
static void launch_uv(sw2_session *s, Ptrs p, int k) {
    k_uv<SCH><<<cells_grid(s->g), cells_block(), 0, s->st>>>(s->g, p, s->fl, s->Dt, s->sigma, s->hv, s->c2, s->c4, s->c6, s->Lu, s->Lv,
                                                             s->LLu, s->LLv, s->LLLu, s->LLLv, s->Du[0], s->Du[1], s->Du[2], s->Dv[0],
                                                             s->Dv[1], s->Dv[2], s->duab, s->dvab, k);
}

#define SCH_SWITCH(F, ...)                 \
    switch (s->sch) {                      \
    case 0: F<0>(__VA_ARGS__); break;      \
    case 1: F<1>(__VA_ARGS__); break;      \
    case 2: F<2>(__VA_ARGS__); break;      \
    case 3: F<3>(__VA_ARGS__); break;      \
    case 4: F<4>(__VA_ARGS__); break;      \
    case 5: F<5>(__VA_ARGS__); break;      \
    case 6: F<6>(__VA_ARGS__); break;      \
    case 7: F<7>(__VA_ARGS__); break;      \
    case 8: F<8>(__VA_ARGS__); break;      \
    default: F<9>(__VA_ARGS__); break;     \
    }

static void lapv(sw2_session *s, const double *uin, const double *vin, double *uo, double *vo) {
    dim3 G = cells_grid(s->g), B = cells_block();
    int r = (s->g.nx + 255) / 256;
    k_lapv_zd<<<G, B, 0, s->st>>>(s->g, s->fl, uin, vin, s->ze, s->di);
    launch_halo(s, s->ze);
    launch_halo(s, s->di);
    k_lapv_fix1<<<r, 256, 0, s->st>>>(s->g, s->fl, uin, s->ze, s->di);
    k_lapv_out<<<G, B, 0, s->st>>>(s->g, s->fl, s->ze, s->di, uo, vo);
    launch_halo(s, uo);
    launch_halo(s, vo);
    k_lapv_fix2<<<r, 256, 0, s->st>>>(s->g, s->fl, uo, vo);
}

static void lape(sw2_session *s, const double *ein, double *eo) {
    dim3 G = cells_grid(s->g), B = cells_block();
    k_lape_px<<<G, B, 0, s->st>>>(s->g, s->fl, ein, s->psi, s->xi);
    k_lape_out<<<G, B, 0, s->st>>>(s->g, s->fl, s->psi, s->xi, eo);
    launch_halo(s, eo);
    k_lape_fix<<<(s->g.nx + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, eo);
}

// storms (and, FP64 only, vortices) of step n at time t: on stream st, into
// the session's fields (meta: the FP32 height anomaly that receives them in
// the mixed form)
static int host_perturbations_on(sw2_session *s, int n, double t, int *eta_changed, hipStream_t st, MField meta) {
    sw *SW = s->SW;
    *eta_changed = 0;
    if (SW->num_Gaussians == 0 && SW->num_vortices == 0) return 0;
    size_t S = (size_t)s->S;
    int K = 0;
    if (SW->num_Gaussians) {
        // the host evaluates each active injection on the cells inside its
        // radius (SW2's expressions and libm) while the device runs this
        // step's kernels; only those cells travel
        K = sw2_gaussians_sparse(SW, s->Dt, t, 8, s->inj);
        if (K < 0) return -1;
        const int b = s->stage_i;
        s->stage_i ^= 1;
        size_t need = 0;
        for (int k = 0; k < K; k++) need += (size_t)s->inj[k].n * (sizeof(int) + 2 * sizeof(double)) + 64;
        if (s->stage_ev[b]) HIPCHECK(hipEventSynchronize(s->stage_ev[b]));
        else HIPCHECK(hipEventCreateWithFlags(&s->stage_ev[b], hipEventDisableTiming));
        if (need > s->stage_cap[b]) {
            if (s->h_stage[b]) HIPCHECK(hipHostFree(s->h_stage[b]));
            s->h_stage[b] = NULL;
            s->stage_cap[b] = 0;
            const size_t cap = need + need / 2 + (1 << 16);
            HIPCHECK(hipHostMalloc((void **)&s->h_stage[b], cap, hipHostMallocDefault));
            s->stage_cap[b] = cap;
        }
        char *stage = s->h_stage[b];
        size_t off = 0;
        for (int k = 0; k < K; k++) {
            const sw2_inj &x = s->inj[k];
            if (x.n == 0) continue;
            int *h_idx = (int *)(stage + off);
            memcpy(h_idx, x.idx, sizeof(int) * (size_t)x.n);
            off += (sizeof(int) * (size_t)x.n + 7) & ~(size_t)7;
            double *h_inc = (double *)(stage + off);
            memcpy(h_inc, x.inc, sizeof(double) * (size_t)x.n);
            off += sizeof(double) * (size_t)x.n;
            double *h_trv = (double *)(stage + off);
            memcpy(h_trv, x.trv, sizeof(double) * (size_t)x.n);
            off += sizeof(double) * (size_t)x.n;
            if (x.n > s->d_cap[k]) {
                HIPCHECK(hipStreamSynchronize(st));
                if (s->d_idx[k]) HIPCHECK(hipFree(s->d_idx[k]));
                if (s->d_inc[k]) HIPCHECK(hipFree(s->d_inc[k]));
                if (s->d_trv[k]) HIPCHECK(hipFree(s->d_trv[k]));
                s->d_idx[k] = NULL;
                s->d_inc[k] = s->d_trv[k] = NULL;
                s->d_cap[k] = 0;
                const int cap = x.n + x.n / 2 + 1024;
                HIPCHECK(hipMalloc((void **)&s->d_idx[k], sizeof(int) * (size_t)cap));
                HIPCHECK(hipMalloc((void **)&s->d_inc[k], sizeof(double) * (size_t)cap));
                HIPCHECK(hipMalloc((void **)&s->d_trv[k], sizeof(double) * (size_t)cap));
                s->d_cap[k] = cap;
            }
            HIPCHECK(hipMemcpyAsync(s->d_idx[k], h_idx, sizeof(int) * (size_t)x.n, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(s->d_inc[k], h_inc, sizeof(double) * (size_t)x.n, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(s->d_trv[k], h_trv, sizeof(double) * (size_t)x.n, hipMemcpyHostToDevice, st));
            if (s->form == SW2_FORM_MIXED)
                k_inject_mixed<<<(x.n + 255) / 256, 256, 0, st>>>(s->g, s->fl, n, x.n, s->d_idx[k], s->d_inc[k], s->d_trv[k],
                                                                    x.has_inc, meta, s->m_tr, s->tracer_active, s->pert);
            else
                k_inject<<<(x.n + 255) / 256, 256, 0, st>>>(s->fl, n, x.n, s->d_idx[k], s->d_inc[k], s->d_trv[k], x.has_inc,
                                                              s->eta, s->tr, s->pert);
            s->perturbs_dirty = 1;
            if (x.has_inc) *eta_changed = 1;
        }
        HIPCHECK(hipEventRecord(s->stage_ev[b], st));
    }
    if (SW->num_vortices) {
        if (!s->h_dv[0][0]) {
            for (int k = 0; k < 8; k++)
                for (int q = 0; q < 3; q++) s->h_dv[q][k] = (double *)malloc(sizeof(double) * S);
            if (dalloc(&s->pinc, S) || dalloc(&s->pdu, S) || dalloc(&s->pdv, S)) return -1;
        }
        int V = sw2_vortices_at(SW, s->Dt, t, 8, s->h_dv[0], s->h_dv[1], s->h_dv[2]);
        if (V < 0) return -1;
        dim3 G = cells_grid(s->g), B = cells_block();
        for (int k = 0; k < V; k++) {
            HIPCHECK(hipMemcpyAsync(s->pinc, s->h_dv[0][k], sizeof(double) * S, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(s->pdu, s->h_dv[1][k], sizeof(double) * S, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(s->pdv, s->h_dv[2][k], sizeof(double) * S, hipMemcpyHostToDevice, st));
            if (s->form == SW2_FORM_MIXED) {  // the halo images follow, once, in step_mixed
                k_add_mixed<<<G, B, 0, st>>>(s->g, s->fl, meta, s->pinc);
                k_add_mixed<<<G, B, 0, st>>>(s->g, s->fl, s->m_u, s->pdu);
                k_add_mixed<<<G, B, 0, st>>>(s->g, s->fl, s->m_v, s->pdv);
            } else {
                k_add<<<G, B, 0, st>>>(s->g, s->fl, s->eta, s->pinc);
                k_add<<<G, B, 0, st>>>(s->g, s->fl, s->u, s->pdu);
                k_add<<<G, B, 0, st>>>(s->g, s->fl, s->v, s->pdv);
                launch_halo(s, s->eta);
                launch_halo(s, s->u);
                launch_halo(s, s->v);
            }
            *eta_changed = 1;
            HIPCHECK(hipStreamSynchronize(st));  // host scratch is reused
        }
    }
    return 0;
}

static int host_perturbations(sw2_session *s, int n, double t, int *eta_changed) {
    return host_perturbations_on(s, n, t, eta_changed, s->st, s->m_eta);
}

template <int SCH, int HV>
static void launch_fused(sw2_session *s, int n, int k, int slot, int late) {
    dim3 G((s->g.nx + FTX - 1) / FTX, (s->g.ny + FTY - 1) / FTY);
    EtaArgs ea{s->eta, s->u, s->v, s->Deta[1], s->Deta[2], s->tr, s->eta2, s->Deta[0], s->Dtr[0], n, k, slot,
               s->tracer_active, s->Dt, s->depth, 2.0 * s->Dt * (s->sigma), 2.0 * s->Dt * (s->tracer_dis), s->c2, s->c4,
               s->courant, s->rest_x, s->rest_y, s->le, s->lle, s->llle, s->c6};
    if (s->kev) (void)hipEventRecord(s->kev[3 * s->kev_i], s->st);
    kf_eta<SCH, HV><<<G, NTE, 0, s->st>>>(s->g, s->fl, ea);
    if (s->kev) (void)hipEventRecord(s->kev[3 * s->kev_i + 1], s->st);
    UVArgs ua{s->eta2, s->u, s->v, s->Du[1], s->Du[2], s->Dv[1], s->Dv[2], s->Dtr[0], s->Dtr[1], s->Dtr[2],
              s->u2, s->v2, s->Du[0], s->Dv[0], s->tr, n, k, slot ^ 1, late, s->Dt, s->sigma, s->c2, s->c4,
              s->Lu, s->Lv, s->LLu, s->LLv, s->LLLu, s->LLLv, s->c6, s->d_cflp};
    kf_uv<SCH, HV><<<G, NTF, 0, s->st>>>(s->g, s->fl, ua);
    if (s->kev) (void)hipEventRecord(s->kev[3 * s->kev_i + 2], s->st);
    km_cfl_parts<<<1, 64, 0, s->st>>>(s->d_cflp, s->fl, slot ^ 1);
}

template <int SCH>
static void launch_fused_hv(sw2_session *s, int n, int k, int slot, int late) {
    if (s->hvmode == 3) launch_fused<SCH, 3>(s, n, k, slot, late);
    else if (s->hvmode) launch_fused<SCH, 2>(s, n, k, slot, late);
    else launch_fused<SCH, 0>(s, n, k, slot, late);
}

// a vortex of SW2's table is added in the step that ends at time t
// (perturbs.c:411-413)
static int vortex_fires(const sw *SW, double Dt, double t) {
    for (int k = 0; k <= SW->num_vortices - 1; k++) {
        const double t0v = SW->vortices[k * N_COLS_VORTICES + 0];
        if ((t0v < t) && (t0v >= (t - Dt))) return 1;
    }
    return 0;
}

// clear the CFL maxima of a slot (before they are reduced again)
__global__ void k_cfl_clear(Flags *fl, int slot) {
    fl->cfl[2 * slot] = 0;
    fl->cfl[2 * slot + 1] = 0;
}

static void swap_ptr(double **a, double **b) {
    double *t = *a;
    *a = *b;
    *b = t;
}

// the fused step: same order and results as step_phased below
static int step_fused(sw2_session *s) {
    const int n = s->n + 1;
    const int k = n - s->nstart - 1;
    const double t = s->t + s->Dt;
    const int slot = n & 1;
    dim3 G = cells_grid(s->g), B = cells_block();
    Ptrs p{s->eta, s->u, s->v, s->wx, s->wy, s->tr};
    if (!s->cfl_ready) {
        k_cfl<<<G, B, 0, s->st>>>(s->g, p, s->fl, slot);
        s->cfl_ready = 1;
    }
    // mode 3: the Laplacians of this step by the per-phase kernels, from the
    // u, v the step starts with (calc.c:324-351, core.c:234-266)
    if (s->hvmode == 3) {
        lape(s, s->u, s->le);
        if (s->lle) {
            lape(s, s->le, s->lle);
            if (s->llle) lape(s, s->lle, s->llle);
        }
        lapv(s, s->u, s->v, s->Lu, s->Lv);
        if (s->LLu) {
            lapv(s, s->Lu, s->Lv, s->LLu, s->LLv);
            if (s->LLLu) lapv(s, s->LLu, s->LLv, s->LLLu, s->LLLv);
        }
    }
    // a vortex added after u and v (core.c:292-296) changes eta, u and v:
    // their checks wait for it, and the next step's CFL maxima are reduced
    // again from the new u and v
    const int vortex = s->SW->num_vortices > 0 && vortex_fires(s->SW, s->Dt, t);
    const int late = ((s->SW->num_Gaussians > 0 || vortex) ? LATE_TE : 0) | (vortex ? LATE_UV : 0);
    // two kernels; the tracer is advanced inside them when it can be
    // nonzero (otherwise once u or v goes non-finite)
    SCH_SWITCH(launch_fused_hv, s, n, k, slot, late);
    swap_ptr(&s->eta, &s->eta2);
    swap_ptr(&s->u, &s->u2);
    swap_ptr(&s->v, &s->v2);
    rot(s->Deta);
    rot(s->Du);
    rot(s->Dv);
    rot(s->Dtr);
    if (s->perturbs_dirty) {
        k_zero_pert<<<(s->S + 255) / 256, 256, 0, s->st>>>(s->fl, n, s->S, s->pert);
        s->perturbs_dirty = 0;
    }
    if (s->SW->num_Gaussians || vortex) {
        // storms and vortices after u and v (core.c:280-300), their halo
        // images, then the checks that the kernels left for after them
        int eta_changed = 0;
        if (host_perturbations(s, n, t, &eta_changed)) return -1;
        launch_halo(s, s->tr, !s->tracer_active);
        if (eta_changed) launch_halo(s, s->eta);
        // a vortex's halo_update of u and v (perturbs.c:433-434) undid the
        // walls the momentum kernel wrote: do_channel again (core.c:303-306)
        if (vortex) k_channel<<<(s->g.nx + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, s->u, s->v);
        k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->tr, n, 2, !s->tracer_active);
        k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->eta, n, 3, 0);
        if (vortex) {
            k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->u, n, 4, 0);
            k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->v, n, 5, 0);
            Ptrs pn{s->eta, s->u, s->v, s->wx, s->wy, s->tr};
            k_cfl_clear<<<1, 1, 0, s->st>>>(s->fl, slot ^ 1);
            k_cfl<<<G, B, 0, s->st>>>(s->g, pn, s->fl, slot ^ 1);
        }
    }
    s->n = n;
    s->t = t;
    return 0;
}

static int step_phased(sw2_session *s);

template <int SCH, int HV, bool TR>
static void launch_mixed(sw2_session *s, int n, int k, int slot, int late) {
    dim3 G((s->g.nx + FTX - 1) / FTX, (s->g.ny + FTY - 1) / FTY);
    MEtaArgs ea{s->m_eta, s->m_u, s->m_v, s->m_De[1], s->m_De[2], s->m_eta2, s->m_De[0], s->d_mt, n, k, slot,
                s->Dt, 2.0 * s->Dt * (s->sigma), s->c2, s->c4, s->courant, s->rest_x, s->rest_y,
                s->m_tr, s->m_Dtr[0], s->tracer_active, 2.0 * s->Dt * (s->tracer_dis), s->le, s->lle, s->llle, s->c6};
    if (s->kev) (void)hipEventRecord(s->kev[3 * s->kev_i], s->st);
    km_eta<SCH, HV, TR><<<G, NTE, 0, s->st>>>(s->g, s->fl, ea);
    if (s->kev) (void)hipEventRecord(s->kev[3 * s->kev_i + 1], s->st);
    MUVArgs ua{s->m_eta2, s->m_u, s->m_v, s->m_Du[1], s->m_Du[2], s->m_Dv[1], s->m_Dv[2], s->m_u2, s->m_v2, s->m_Du[0],
               s->m_Dv[0], s->d_mt, s->d_cflp, n, k, s->Dt, s->sigma, s->c2, s->c4, late,
               s->m_tr, s->m_Dtr[0], s->m_Dtr[1], s->m_Dtr[2], s->tracer_active,
               s->Lu, s->Lv, s->LLu, s->LLv, s->LLLu, s->LLLv, s->c6};
    km_uv<SCH, HV, TR><<<G, NTU, 0, s->st>>>(s->g, s->fl, ua);
    if (s->kev) (void)hipEventRecord(s->kev[3 * s->kev_i + 2], s->st);
    km_cfl_parts<<<1, 64, 0, s->st>>>(s->d_cflp, s->fl, slot ^ 1);
}

// the kernels with the tracer's code only for a world whose tracer runs
template <int SCH>
static void launch_mixed_hv(sw2_session *s, int n, int k, int slot, int late) {
    if (s->tracer_active) {
        if (s->mhv == 3) launch_mixed<SCH, 3, true>(s, n, k, slot, late);
        else if (s->mhv) launch_mixed<SCH, 2, true>(s, n, k, slot, late);
        else launch_mixed<SCH, 0, true>(s, n, k, slot, late);
    } else {
        if (s->mhv == 3) launch_mixed<SCH, 3, false>(s, n, k, slot, late);
        else if (s->mhv) launch_mixed<SCH, 2, false>(s, n, k, slot, late);
        else launch_mixed<SCH, 0, false>(s, n, k, slot, late);
    }
}

static int step_mixed(sw2_session *s) {
    const int n = s->n + 1;
    const int k = n - s->nstart - 1;
    const int slot = n & 1;
    if (!s->cfl_ready) {
        km_cfl<<<1024, 256, 0, s->st>>>(s->g, s->fl, s->m_u, s->m_v, slot);
        s->cfl_ready = 1;
    }
    const double t = s->t + s->Dt;
    // a vortex added after u and v (core.c:292-296) changes eta, u and v:
    // their checks and the wall wait for it, and the next step's CFL maxima
    // are reduced again from the new u and v (as on the fused FP64 step)
    const int vortex = s->SW->num_vortices > 0 && vortex_fires(s->SW, s->Dt, t);
    const int late = ((s->SW->num_Gaussians > 0 || vortex) ? MLATE_TE : 0) | (vortex ? MLATE_UV : 0);
    if (s->mhv == 3) {  // del-6: SW2's three passes of each Laplacian (calc.c:324-351, core.c:234-266)
        k_mixed_f64<<<(s->S + 255) / 256, 256, 0, s->st>>>(s->g, s->m_u, s->m_u64, s->S);
        k_mixed_f64<<<(s->S + 255) / 256, 256, 0, s->st>>>(s->g, s->m_v, s->m_v64, s->S);
        lape(s, s->m_u64, s->le);
        lape(s, s->le, s->lle);
        lape(s, s->lle, s->llle);
        lapv(s, s->m_u64, s->m_v64, s->Lu, s->Lv);
        lapv(s, s->Lu, s->Lv, s->LLu, s->LLv);
        lapv(s, s->LLu, s->LLv, s->LLLu, s->LLLv);
    }
    SCH_SWITCH(launch_mixed_hv, s, n, k, slot, late);
    std::swap(s->m_eta, s->m_eta2);
    std::swap(s->m_u, s->m_u2);
    std::swap(s->m_v, s->m_v2);
    frot(s->m_De);
    frot(s->m_Du);
    frot(s->m_Dv);
    if (s->tracer_active) frot(s->m_Dtr);
    if (s->SW->num_Gaussians || vortex) {
        // storms and vortices after u and v, then the checks (core.c:280-312)
        if (s->perturbs_dirty && s->pert) {
            k_zero_pert<<<(s->S + 255) / 256, 256, 0, s->st>>>(s->fl, n, s->S, s->pert);
            s->perturbs_dirty = 0;
        }
        int eta_changed = 0;
        if (host_perturbations(s, n, t, &eta_changed)) return -1;
        const dim3 G = cells_grid(s->g), B = cells_block();
        if (vortex) {  // the vortices' halo_update of eta, u, v, then do_channel (perturbs.c:433-434, core.c:298-306)
            const int cnt = 4 * (s->g.nx + 4) + 4 * s->g.ny;
            k_halo_mixed<<<(cnt + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, s->m_eta.hi);
            k_halo_mixed<<<(cnt + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, s->m_u.hi);
            k_halo_mixed<<<(cnt + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, s->m_v.hi);
            k_channel_mixed<<<(s->g.nx + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, s->m_u, s->m_v);
        }
        if (s->tracer_active) k_check_mixed<<<G, B, 0, s->st>>>(s->g, s->fl, s->m_tr, nullptr, n, 2);
        k_check_mixed<<<G, B, 0, s->st>>>(s->g, s->fl, s->m_eta, s->d_mt + MT_HBAR * s->g.R, n, 3);
        if (vortex) {
            k_check_mixed<<<G, B, 0, s->st>>>(s->g, s->fl, s->m_u, nullptr, n, 4);
            k_check_mixed<<<G, B, 0, s->st>>>(s->g, s->fl, s->m_v, nullptr, n, 5);
            k_cfl_clear<<<1, 1, 0, s->st>>>(s->fl, slot ^ 1);
            km_cfl<<<1024, 256, 0, s->st>>>(s->g, s->fl, s->m_u, s->m_v, slot ^ 1);
        }
    }
    s->n = n;
    s->t = t;
    return 0;
}

static int step_once(sw2_session *s) {
    if (s->form == SW2_FORM_MIXED) return step_mixed(s);
    return s->fused ? step_fused(s) : step_phased(s);
}

static int step_phased(sw2_session *s) {
    const int n = s->n + 1;
    const int k = n - s->nstart - 1;
    const double t = s->t + s->Dt;
    const int slot = n & 1;
    dim3 G = cells_grid(s->g), B = cells_block();
    Ptrs p{s->eta, s->u, s->v, s->wx, s->wy, s->tr};
    k_cfl<<<G, B, 0, s->st>>>(s->g, p, s->fl, slot);
    k_cfl_fin<<<1, 1, 0, s->st>>>(s->fl, slot, n, s->courant, s->Dt, s->rest_x, s->rest_y);
    {  // tracer: skipped on the device while it is identically zero and u, v finite
        const int gated = !s->tracer_active;
        double sig2 = 2.0 * s->Dt * (s->tracer_dis);
        SCH_SWITCH(launch_flux, s, p, true, s->tr, sig2, s->Dtr[0]);
        k_tracer_update<<<G, B, 0, s->st>>>(s->g, s->fl, s->tr, s->Dtr[0], s->Dtr[1], s->Dtr[2], k, gated);
        rot(s->Dtr);
        launch_halo(s, s->tr, gated);
    }
    double sig2 = 2.0 * s->Dt * (s->sigma);
    SCH_SWITCH(launch_flux, s, p, false, s->eta, sig2, s->Deta[0]);
    if (s->hv) {
        lape(s, s->u, s->le);
        if (s->lle) {
            lape(s, s->le, s->lle);
            if (s->llle) lape(s, s->lle, s->llle);
        }
    }
    k_eta_update<<<G, B, 0, s->st>>>(s->g, p, s->fl, s->Deta[0], s->Deta[1], s->Deta[2], s->le, s->lle, s->llle, s->hv, s->c2,
                                      s->c4, s->c6, k, s->Dt, s->depth);
    rot(s->Deta);
    launch_halo(s, s->eta);
    if (s->hv) {
        lapv(s, s->u, s->v, s->Lu, s->Lv);
        if (s->LLu) {
            lapv(s, s->Lu, s->Lv, s->LLu, s->LLv);
            if (s->LLLu) lapv(s, s->LLu, s->LLv, s->LLLu, s->LLLv);
        }
    }
    SCH_SWITCH(launch_uv, s, p, k);
    rot(s->Du);
    rot(s->Dv);
    k_cor_u<<<G, B, 0, s->st>>>(s->g, p, s->fl, s->duab, s->Dt);
    k_cor_v<<<G, B, 0, s->st>>>(s->g, p, s->fl, s->dvab, s->Dt);
    launch_halo(s, s->u);
    launch_halo(s, s->v);
    if (s->perturbs_dirty) {
        k_zero_pert<<<(s->S + 255) / 256, 256, 0, s->st>>>(s->fl, n, s->S, s->pert);
        s->perturbs_dirty = 0;
    }
    int eta_changed = 0;
    if (host_perturbations(s, n, t, &eta_changed)) return -1;
    launch_halo(s, s->tr, !s->tracer_active);
    if (eta_changed) launch_halo(s, s->eta);
    k_channel<<<(s->g.nx + 255) / 256, 256, 0, s->st>>>(s->g, s->fl, s->u, s->v);
    k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->tr, n, 2, !s->tracer_active);
    k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->eta, n, 3, 0);
    k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->u, n, 4, 0);
    k_check<<<G, B, 0, s->st>>>(s->g, s->fl, s->v, n, 5, 0);
    s->n = n;
    s->t = t;
    return 0;
}

extern "C" int sw2_session_step(sw2_session *s, int nsteps) {
    for (int q = 0; q < nsteps; q++) {
        if (step_once(s)) return -1;
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->st));
    return 0;
}

extern "C" int sw2_session_time(sw2_session *s, int nsteps, double *ms) {
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, s->st));
    for (int q = 0; q < nsteps; q++)
        if (step_once(s)) return -1;
    HIPCHECK(hipEventRecord(e1, s->st));
    HIPCHECK(hipEventSynchronize(e1));
    float f = 0;
    HIPCHECK(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    return 0;
}

extern "C" int sw2_session_kernel_times(sw2_session *s, int nsteps, double *ms, int nk) {
    if (!s->fused) {
        sw2_set_error("per-kernel times are measured on the fused path only");
        return -1;
    }
    if (nk < 2 || nsteps < 1) {
        sw2_set_error("need nk >= 2 and nsteps >= 1");
        return -1;
    }
    hipEvent_t *ev = (hipEvent_t *)calloc((size_t)3 * nsteps, sizeof(hipEvent_t));
    int rc = 0;
    for (int q = 0; q < 3 * nsteps && !rc; q++)
        if (hipEventCreate(&ev[q]) != hipSuccess) rc = -1;
    s->kev = ev;
    for (int q = 0; q < nsteps && !rc; q++) {
        s->kev_i = q;
        if (step_once(s)) rc = -1;
    }
    s->kev = NULL;
    double te = 0, tu = 0;
    if (!rc && hipEventSynchronize(ev[3 * nsteps - 1]) != hipSuccess) rc = -1;
    for (int q = 0; q < nsteps && !rc; q++) {
        float a = 0, b = 0;
        if (hipEventElapsedTime(&a, ev[3 * q], ev[3 * q + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[3 * q + 1], ev[3 * q + 2]) != hipSuccess)
            rc = -1;
        te += a;
        tu += b;
    }
    for (int q = 0; q < 3 * nsteps; q++)
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    free(ev);
    if (rc) {
        if (!sw2_last_error()[0]) sw2_set_error("event timing failed");
        return -1;
    }
    ms[0] = te;
    ms[1] = tu;
    return 0;
}

extern "C" int sw2_session_energy(sw2_session *s, double *ke, double *ape, double *mass) {
    const int nb = ENERGY_PARTS;
    double *out = s->d_part + 3 * nb;
    if (s->form == SW2_FORM_MIXED)
        k_energy<float><<<nb, 256, 0, s->st>>>(s->g, s->m_eta.hi, s->m_u.hi, s->m_v.hi, s->d_mt + MT_HBAR * s->g.R, s->depth,
                                               s->d_part, s->m_eta.lo, s->m_u.lo, s->m_v.lo);
    else
        k_energy<double><<<nb, 256, 0, s->st>>>(s->g, s->eta, s->u, s->v, nullptr, s->depth, s->d_part);
    k_energy_fin<<<1, 64, 0, s->st>>>(nb, s->d_part, out);
    double h[3];
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, s->st));
    HIPCHECK(hipStreamSynchronize(s->st));
    if (ke) *ke = h[0];
    if (ape) *ape = h[1];
    if (mass) *mass = h[2];
    return 0;
}

extern "C" int sw2_session_energy_noalloc(sw2_session *s, double *d_out3, void *stream) {
    if (!s || !d_out3) {
        sw2_set_error("sw2_session_energy_noalloc: session and device output needed");
        return -1;
    }
    const int nb = ENERGY_PARTS;  // the same reduction tree as sw2_session_energy
    double *part = nullptr;
    if (scratch_address(&part)) return -1;
    hipStream_t st = stream ? (hipStream_t)stream : s->st;
    // every block writes its own partials and k_energy_fin writes all three
    // outputs, so nothing in the scratch has to start at zero
    if (s->form == SW2_FORM_MIXED)
        k_energy<float><<<nb, 256, 0, st>>>(s->g, s->m_eta.hi, s->m_u.hi, s->m_v.hi, s->d_mt + MT_HBAR * s->g.R, s->depth, part,
                                            s->m_eta.lo, s->m_u.lo, s->m_v.lo);
    else
        k_energy<double><<<nb, 256, 0, st>>>(s->g, s->eta, s->u, s->v, nullptr, s->depth, part);
    k_energy_fin<<<1, 64, 0, st>>>(nb, part, d_out3);
    HIPCHECK(hipGetLastError());
    return 0;
}

extern "C" int sw2_session_download(sw2_session *s, sw *SW) {
    size_t S = (size_t)s->S;
    HIPCHECK(hipStreamSynchronize(s->st));
    if (s->form == SW2_FORM_MIXED) {
        float *h = (float *)malloc(sizeof(float) * S), *hl = (float *)malloc(sizeof(float) * S);
        MField src[4] = {s->m_eta, s->m_u, s->m_v, s->m_tr};
        double *dst[4] = {SW->eta, SW->u, SW->v, SW->tracer};
        int rc = 0;
        for (int f = 0; f < (s->tracer_active ? 4 : 3) && !rc; f++) {
            if (hipMemcpy(h, src[f].hi, sizeof(float) * S, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(hl, src[f].lo, sizeof(float) * S, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = -1;
                break;
            }
            // interior cells: value + residual; halo cells: the value of
            // their image, completed by the residual of the interior cell
            // it images (the walls' copies and zeros as stored)
            for (int j = -1; j <= SW->ny + 2; j++)
                for (int i = -1; i <= SW->nx + 2; i++) {
                    int q = LP(i, j, SW->M);
                    double x = (double)h[q];
                    if (j >= 1 && j <= SW->ny) {
                        const int ii = i < 1 ? i + SW->nx : (i > SW->nx ? i - SW->nx : i);
                        x += (double)hl[LP(ii, j, SW->M)];
                    }
                    dst[f][q] = x + (f == 0 ? s->h_hbar[j + 1] : 0.0);
                }
        }
        free(h);
        free(hl);
        if (!rc && s->pert && hipMemcpy(SW->perturbs, s->pert, sizeof(double) * S, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
        if (rc) sw2_set_error("download failed");
        return rc;
    }
    struct {
        double *d;
        double *h;
    } dn[] = {{s->eta, SW->eta}, {s->u, SW->u}, {s->v, SW->v}, {s->tr, SW->tracer}, {s->pert, SW->perturbs}};
    for (auto &x : dn) HIPCHECK(hipMemcpy(x.h, x.d, sizeof(double) * S, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int sw2_session_failed_step(sw2_session *s, char *msg, int msglen) {
    Flags f;
    if (hipMemcpy(&f, s->fl, sizeof f, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (!f.fail_step) return 0;
    if (msg && msglen > 0) {
        static const char *names[] = {"", "cfl", "tracer", "eta", "u", "v"};
        int kind = 0;
        for (int b = 2; b <= 5 && !kind; b++)
            if (f.kind_mask & (1 << b)) kind = b;
        if (f.fail_kind == 1)
            snprintf(msg, msglen, "Smallest possible timestep is Dt_CFL=%.6e while Dt=%.6e.\n", f.fail_val, s->Dt);
        else
            snprintf(msg, msglen, "OOOPS nonsense value detected in field <%s> at step %d; range acceptable = %e to %e \n",
                     names[kind], f.fail_step, -1e10, 1e10);
    }
    return f.fail_step;
}

extern "C" int sw2_session_last_cfl(sw2_session *s, double *dtcfl) {
    Flags f;
    HIPCHECK(hipStreamSynchronize(s->st));
    HIPCHECK(hipMemcpy(&f, s->fl, sizeof f, hipMemcpyDeviceToHost));
    *dtcfl = f.last_dtcfl;
    return 0;
}

extern "C" int sw2_session_state(sw2_session *s, int *n, double *t) {
    if (n) *n = s->n;
    if (t) *t = s->t;
    return 0;
}

extern "C" void sw2_session_destroy(sw2_session *s) {
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    double *all[] = {s->d_tab, s->eta, s->u, s->v, s->eta2, s->u2, s->v2, s->tr, s->pert, s->wx, s->wy, s->duab, s->dvab, s->Lu, s->Lv, s->LLu, s->LLv,
                     s->LLLu, s->LLLv, s->ze, s->di, s->le, s->lle, s->llle, s->psi, s->xi, s->pinc, s->pdu, s->pdv};
    for (auto p : all)
        if (p) (void)hipFree(p);
    for (int k = 0; k < 3; k++) {
        if (s->Deta[k]) (void)hipFree(s->Deta[k]);
        if (s->Du[k]) (void)hipFree(s->Du[k]);
        if (s->Dv[k]) (void)hipFree(s->Dv[k]);
        if (s->Dtr[k]) (void)hipFree(s->Dtr[k]);
    }
    MField fall[] = {s->m_eta, s->m_u, s->m_v, s->m_eta2, s->m_u2, s->m_v2, s->m_tr};
    for (auto p : fall) {
        if (p.hi) (void)hipFree(p.hi);
        if (p.lo) (void)hipFree(p.lo);
    }
    for (int k = 0; k < 3; k++) {
        if (s->m_De[k]) (void)hipFree(s->m_De[k]);
        if (s->m_Du[k]) (void)hipFree(s->m_Du[k]);
        if (s->m_Dv[k]) (void)hipFree(s->m_Dv[k]);
        if (s->m_Dtr[k]) (void)hipFree(s->m_Dtr[k]);
    }
    if (s->m_u64) (void)hipFree(s->m_u64);
    if (s->m_v64) (void)hipFree(s->m_v64);
    if (s->d_mt) (void)hipFree(s->d_mt);
    if (s->d_part) (void)hipFree(s->d_part);
    if (s->d_cflp) (void)hipFree(s->d_cflp);
    free(s->h_hbar);
    if (s->fl && !s->fl_borrowed) (void)hipFree(s->fl);
    if (s->st) (void)hipStreamDestroy(s->st);
    for (int k = 0; k < 8; k++) {
        for (int q = 0; q < 3; q++) free(s->h_dv[q][k]);
        if (s->d_idx[k]) (void)hipFree(s->d_idx[k]);
        if (s->d_inc[k]) (void)hipFree(s->d_inc[k]);
        if (s->d_trv[k]) (void)hipFree(s->d_trv[k]);
    }
    sw2_inj_free(s->inj, 8);
    for (int b = 0; b < 2; b++) {
        if (s->h_stage[b]) (void)hipHostFree(s->h_stage[b]);
        if (s->stage_ev[b]) (void)hipEventDestroy(s->stage_ev[b]);
    }
    free(s);
}

extern "C" const char *sw2_build_info(void) {
    return "libsw2 gfx950 FP64 -O3 -ffp-contract=off";
}

extern "C" int sw2_session_fused(sw2_session *s) { return s->fused; }

extern "C" int sw2_session_form(sw2_session *s) { return s->form; }

// ====================================================================== batches
// A batch owns up to `cap` member slots.  Each member is a mixed-form session
// (its own FP32 state, tables and guard record); the batch steps all of them
// together with the persistent kernels of sw2_mixed.inc, from a graph of
// `gsteps` steps captured once.  Members join and leave between steps: a
// join writes the member's record and zeroes its step count, a leave empties
// its slot, and neither changes the launches.
struct sw2_batch {
    int form, cap, sch, hvmode, nmem;
    size_t budget, used;
    sw2_session **sess;
    size_t *foot;
    BMember *h_mem;
    BHdr h_hdr;  // the kernels' header, passed by value
    BMember *d_mem;
    BState *d_st;
    Flags *d_flags;  // the members' guard records, by slot
    hipStream_t st;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    int gsteps, blocks_eta, blocks_uv, table_dirty;
    int storms;   // members with in-loop storms (their storms run after each step)
    int tracers;  // members whose tracer runs: the kernels with the tracer's code step the batch
    int graph_tr; // whether the captured graph holds those kernels
    int graph_stale;  // the tile count changed since the capture (one workgroup per tile)
    // storms in a replayed graph: the graph's storm kernels cover graph_nl
    // members (the storm members when it was captured) and read the window
    // table d_wt (gsteps x graph_nl entries) and the cell lists in d_cells,
    // which the host fills ahead of each replay through pinned staging
    // (two sets, used alternately, each reused once its copies have run)
    int graph_nl, wt_cap;
    SEntry *d_wt, *h_wt;
    char *d_cells, *h_build;
    size_t cells_cap, build_cap;
    char *h_stage[2];
    size_t stage_cap[2];
    hipEvent_t stage_ev[2];
    int stage_i;
    int *paused;  // host copy of the members' paused flags
    int2 *h_tmap, *d_tmap;  // every flat tile's member slot and tile coordinates, host and device
    int tmap_cap;           // entries of both
    int xshift_set;         // log2 of the XCD block size when set (SW2_XCD_BLOCK), else -1
    int shape;             // tile shape of the batch kernels: an index of BATCH_SHAPES
};

// tile shapes of the batch kernels (columns x rows), in the order of the
// namespaces that hold their kernels
struct BatchShape {
    int tx, ty;
};
static const BatchShape BATCH_SHAPES[] = {{32, 16}, {64, 8}, {64, 16}};
constexpr int NBATCH_SHAPES = 3;
static_assert(FTX == 32 && FTY == 16, "the single-run kernels use the first batch shape");

extern "C" size_t sw2_batch_footprint(const sw *SW, int form) {
    const size_t S = (size_t)(SW->nx + 4) * (SW->ny + 4), R = (size_t)SW->ny + 4;
    size_t f = (size_t)(SW2_NT + 2) * R * sizeof(double) + 3 * R * sizeof(double) + (3 * 1024 + 3) * sizeof(double) + sizeof(Flags);
    int tracer = 0;  /* the tracer can be nonzero: a storm injects it, or it is nonzero already */
    for (int k = 0; k < SW->num_Gaussians; k++)
        if (SW->Gaussians[k * N_COLS_GAUSS + 11] != 0) tracer = 1;
    if (!tracer)
        for (size_t q = 0; q < S && !tracer; q++) tracer = SW->tracer[q] != 0;
    if (form == SW2_FORM_MIXED) f += (12 + 9 + (tracer ? 5 : 0)) * S * sizeof(float);
    else f += 24 * S * sizeof(double);
    if (SW->num_Gaussians) f += S * sizeof(double); /* the storms' marks */
    return f;
}

extern "C" int sw2_device_cus(void) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || ncu < 1) ncu = 256;
    return ncu;
}

extern "C" int sw2_batch_create(int form, int capacity, double hbm_budget_bytes, sw2_batch **out) {
    *out = NULL;
    if (form != SW2_FORM_MIXED) {
        sw2_set_error("batches step the mixed-precision form (SW2_FORM_MIXED)");
        return -1;
    }
    if (capacity < 1 || capacity > 4096) {
        sw2_set_error("capacity %d out of 1..4096", capacity);
        return -1;
    }
    sw2_batch *b = (sw2_batch *)calloc(1, sizeof(sw2_batch));
    b->form = form;
    b->cap = capacity;
    b->sch = -1;
    b->budget = hbm_budget_bytes > 0 ? (size_t)hbm_budget_bytes : (size_t)-1;
    b->sess = (sw2_session **)calloc(capacity, sizeof(sw2_session *));
    b->foot = (size_t *)calloc(capacity, sizeof(size_t));
    b->h_mem = (BMember *)calloc(capacity, sizeof(BMember));
    b->paused = (int *)calloc(capacity, sizeof(int));
    b->gsteps = 8;
    // default grids: one workgroup per tile, so the hardware balances the
    // tiles over the compute units as they free up (on the 64-member
    // ensemble 7% faster than as many persistent workgroups as stay
    // resident, whose static stride cannot balance uneven tiles)
    b->blocks_eta = 0;
    b->blocks_uv = 0;
    int rc = 0;
    if (hipMalloc((void **)&b->d_mem, sizeof(BMember) * capacity) != hipSuccess) rc = -1;
    if (!rc && hipMalloc((void **)&b->d_st, sizeof(BState) * capacity) != hipSuccess) rc = -1;
    if (!rc && hipMalloc((void **)&b->d_flags, sizeof(Flags) * capacity) != hipSuccess) rc = -1;
    if (!rc && hipMemset(b->d_flags, 0, sizeof(Flags) * capacity) != hipSuccess) rc = -1;
    if (!rc && hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) rc = -1;
    if (!rc) {
        BState *z = (BState *)calloc(capacity, sizeof(BState));
        for (int m = 0; m < capacity; m++) z[m].done = 1;
        if (hipMemcpy(b->d_st, z, sizeof(BState) * capacity, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
        free(z);
    }
    b->h_hdr.nslots = capacity;
    b->h_hdr.flags = b->d_flags;
    b->h_hdr.total_tiles = 0;
    b->xshift_set = -1;
    if (const char *e = getenv("SW2_XCD_BLOCK")) {  // a power of two, 8 .. 4096
        const int v = atoi(e);
        for (int k = 3; k <= 12; k++)
            if (v == 1 << k) b->xshift_set = k;
    }
    b->table_dirty = 1;
    if (rc) {
        sw2_set_error("batch allocation failed");
        sw2_batch_destroy(b);
        return -1;
    }
    *out = b;
    return 0;
}

// number the tiles of the members that advance: a held member keeps its
// slot but gets no tiles (an empty span at its tile0, which the kernels'
// slot search never returns), so a launch with one workgroup per tile has
// no workgroups for held runs
static void batch_retile(sw2_batch *b) {
    int t = 0;
    for (int m = 0; m < b->cap; m++) {
        b->h_mem[m].tile0 = t;
        if (!b->paused[m]) t += b->h_mem[m].ntx * b->h_mem[m].nty;
    }
    // the graph holds the old tile count (in its header, and in its grids
    // when launching one workgroup per tile): it is captured again at the
    // next step (a replay may still be running, so it is not destroyed here)
    if (t != b->h_hdr.total_tiles && b->gexec) b->graph_stale = 1;
    b->h_hdr.total_tiles = t;
    // XCD blocks of 256 tiles, fewer when the tiles make less than two
    // rounds of eight such blocks, so that small batches are dealt too
    // (measured, a lone 1024 x 512 member: 1.22 -> 1.10 x its solo step
    // with 64-tile blocks; 64 to 2048 within 2% on large ones)
    int xs = 8;
    while (xs > 3 && t < (16 << xs)) xs--;
    b->h_hdr.xshift = b->xshift_set >= 0 ? b->xshift_set : xs;
    b->table_dirty = 1;
}

extern "C" int sw2_batch_add(sw2_batch *b, sw *SW, int nsteps, int *id) {
    if (nsteps < 1) {
        sw2_set_error("a member needs at least one step");
        return -1;
    }
    int slot = -1;
    for (int m = 0; m < b->cap && slot < 0; m++)
        if (!b->sess[m]) slot = m;
    if (slot < 0) {
        sw2_set_error("batch full (%d members)", b->cap);
        return 1;
    }
    const size_t foot = sw2_batch_footprint(SW, b->form);
    if (b->used + foot > b->budget) {
        sw2_set_error("HBM budget: %zu bytes in use, member needs %zu, budget %zu", b->used, foot, b->budget);
        return 1;
    }
    if (SW->num_vortices || SW->nu6 != 0) {  // a vortex step's late checks and wall, del-6's per-phase
        // Laplacians: single sessions only
        sw2_set_error("batches take cases without vortices or del-6 (a mixed session runs them)");
        return -1;
    }
    sw2_session *s = NULL;
    // the session's guard record is the slot's in the batch's array
    if (session_create(SW, b->form, &s, b->d_flags + slot)) return -1;
    if (b->sch < 0) {
        b->sch = s->sch;
        b->hvmode = s->hvmode;
    } else if (s->sch != b->sch || s->hvmode != b->hvmode) {
        sw2_session_destroy(s);
        sw2_set_error("members of a batch share the limiter and hyperviscosity order");
        return -1;
    }
    BMember &M = b->h_mem[slot];
    memset(&M, 0, sizeof M);
    M.g = s->g;
    if (!s->d_cflp) {
        HIPCHECK(hipMalloc((void **)&s->d_cflp, sizeof(unsigned long long) * 2 * NPART));
        HIPCHECK(hipMemset(s->d_cflp, 0, sizeof(unsigned long long) * 2 * NPART));
    }
    M.cflp = s->d_cflp;
    M.mt = s->d_mt;
    M.eta[0] = s->m_eta;
    M.eta[1] = s->m_eta2;
    M.u[0] = s->m_u;
    M.u[1] = s->m_u2;
    M.v[0] = s->m_v;
    M.v[1] = s->m_v2;
    for (int k = 0; k < 3; k++) {
        M.De[k] = s->m_De[k];
        M.Du[k] = s->m_Du[k];
        M.Dv[k] = s->m_Dv[k];
        M.Dtr[k] = s->m_Dtr[k];
    }
    M.ntx = (SW->nx + BATCH_SHAPES[b->shape].tx - 1) / BATCH_SHAPES[b->shape].tx;
    M.nty = (SW->ny + BATCH_SHAPES[b->shape].ty - 1) / BATCH_SHAPES[b->shape].ty;
    if (M.ntx > 0xffff || M.nty > 0x7fff) {  // the tile map packs a tile's column and row into 16 + 15 bits
        sw2_session_destroy(s);
        memset(&M, 0, sizeof M);
        sw2_set_error("a member of %d x %d cells has more tiles per row or column than a batch numbers", SW->nx, SW->ny);
        return -1;
    }
    M.nstart = 0;
    M.nlast = nsteps;
    M.Dt = s->Dt;
    M.sig2 = 2.0 * s->Dt * (s->sigma);
    M.sigma = s->sigma;
    M.c2 = s->c2;
    M.c4 = s->c4;
    M.courant = s->courant;
    M.rest_x = s->rest_x;
    M.rest_y = s->rest_y;
    M.late = SW->num_Gaussians > 0;
    M.tr = s->m_tr;
    M.tr_on = s->tracer_active;
    M.tsig2 = 2.0 * s->Dt * (s->tracer_dis);
    if (M.late) b->storms++;
    if (M.tr_on) b->tracers++;
    b->paused[slot] = 0;
    b->sess[slot] = s;
    b->foot[slot] = foot;
    b->used += foot;
    b->nmem++;
    batch_retile(b);
    // a fresh guard record, the CFL maxima of the member's first step and
    // its step state, queued behind whatever the batch stream still runs:
    // admission does not wait for the device (the session's uploads above
    // have completed), and the member's first step is the batch's next
    HIPCHECK(hipMemsetAsync(s->fl, 0, sizeof(Flags), b->st));
    km_cfl<<<256, 256, 0, b->st>>>(s->g, s->fl, s->m_u, s->m_v, 1);
    kb_join<<<1, 64, 0, b->st>>>(b->d_st, slot);
    HIPCHECK(hipGetLastError());
    *id = slot;
    return 0;
}

static void batch_launch_step(sw2_batch *b) {
    const int tiles = b->h_hdr.total_tiles > 0 ? b->h_hdr.total_tiles : 1;
    const int ge = b->blocks_eta ? b->blocks_eta : tiles, gu = b->blocks_uv ? b->blocks_uv : tiles;
    // the kernels with the tracer's code only while a member's tracer runs;
    // persistent workgroups exist for the 32 x 16 tiles only
    const int hv = b->hvmode ? 2 : 0, tr = b->tracers > 0;
    if (b->shape == 1)
        sw2_batch_phases_64x8(b->d_mem, b->d_st, &b->h_hdr, b->sch, hv, tr, ge, gu, 0, 0, b->st);
    else if (b->shape == 2)
        sw2_batch_phases_64x16(b->d_mem, b->d_st, &b->h_hdr, b->sch, hv, tr, ge, gu, 0, 0, b->st);
    else
        sw2_batch_phases_32x16(b->d_mem, b->d_st, &b->h_hdr, b->sch, hv, tr, ge, gu, b->blocks_eta != 0,
                               b->blocks_uv != 0, b->st);
    kb_adv<<<(b->cap + 3) / 4, 256, 0, b->st>>>(b->d_mem, b->d_st, b->h_hdr);
}

static int batch_upload(sw2_batch *b) {
    if (!b->table_dirty) return 0;
    // the tile map: flat tile -> member slot, in the numbering batch_retile
    // made (held members and empty slots own no tiles).  It grows by
    // doubling; the kernels take its address in the header (a new address
    // makes the graph stale).
    const int t = b->h_hdr.total_tiles;
    if (t > b->tmap_cap) {
        const int cap = t > 2 * b->tmap_cap ? t : 2 * b->tmap_cap;
        int2 *nh = (int2 *)realloc(b->h_tmap, sizeof(int2) * (size_t)cap);
        if (!nh) {
            sw2_set_error("out of host memory for the tile map of %d tiles", cap);
            return -1;
        }
        b->h_tmap = nh;
        HIPCHECK(hipStreamSynchronize(b->st));
        if (b->d_tmap) HIPCHECK(hipFree(b->d_tmap));
        b->d_tmap = NULL;
        HIPCHECK(hipMalloc((void **)&b->d_tmap, sizeof(int2) * (size_t)cap));
        b->tmap_cap = cap;
    }
    for (int m = 0, k = 0; m < b->cap; m++)
        if (!b->paused[m])
            for (int ty = 0; ty < b->h_mem[m].nty; ty++)
                for (int tx = 0; tx < b->h_mem[m].ntx; tx++) b->h_tmap[k++] = make_int2(m, tx | (ty << 16));
    if (b->h_hdr.tmap != b->d_tmap && b->gexec) b->graph_stale = 1;
    b->h_hdr.tmap = b->d_tmap;
    if (t > 0) HIPCHECK(hipMemcpyAsync(b->d_tmap, b->h_tmap, sizeof(int2) * (size_t)t, hipMemcpyHostToDevice, b->st));
    HIPCHECK(hipMemcpyAsync(b->d_mem, b->h_mem, sizeof(BMember) * b->cap, hipMemcpyHostToDevice, b->st));
    HIPCHECK(hipStreamSynchronize(b->st));
    b->table_dirty = 0;
    return 0;
}

static void batch_drop_graph(sw2_batch *b) {
    b->graph_stale = 0;
    if (!b->gexec) return;
    (void)hipStreamSynchronize(b->st);
    (void)hipGraphExecDestroy(b->gexec);
    (void)hipGraphDestroy(b->graph);
    b->gexec = NULL;
    b->graph = NULL;
}

// the storm kernels of window step q, for the graph_nl storm members
static void batch_launch_storms(sw2_batch *b, int q) {
    const int nl = b->graph_nl;
    const SEntry *w = b->d_wt + (size_t)q * nl;
    kbs_zero<<<dim3(256, nl), 256, 0, b->st>>>(w);
    kbs_inject<<<dim3(1, nl), 1024, 0, b->st>>>(w);
    kbs_check<<<dim3(256, nl), 256, 0, b->st>>>(w);
}

static int batch_capture(sw2_batch *b) {
    const int nl = b->storms;
    if (b->gexec && !b->graph_stale && b->graph_tr == (b->tracers > 0) && b->graph_nl == nl) return 0;
    // captured for another tile count, with the other kernel variant (tracer
    // code in or out) or for another number of storm members
    batch_drop_graph(b);
    if (nl && b->gsteps * nl > b->wt_cap) {
        const int cap = b->gsteps * nl;
        SEntry *h = (SEntry *)realloc(b->h_wt, sizeof(SEntry) * (size_t)cap);
        if (!h) {
            sw2_set_error("out of host memory for the storm window table");
            return -1;
        }
        b->h_wt = h;
        HIPCHECK(hipStreamSynchronize(b->st));
        if (b->d_wt) HIPCHECK(hipFree(b->d_wt));
        b->d_wt = NULL;
        b->wt_cap = 0;
        HIPCHECK(hipMalloc((void **)&b->d_wt, sizeof(SEntry) * (size_t)cap));
        b->wt_cap = cap;
    }
    b->graph_tr = b->tracers > 0;
    b->graph_nl = nl;
    HIPCHECK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < b->gsteps; q++) {
        batch_launch_step(b);
        if (nl) batch_launch_storms(b, q);
    }
    HIPCHECK(hipStreamEndCapture(b->st, &b->graph));
    HIPCHECK(hipGraphInstantiate(&b->gexec, b->graph, NULL, NULL, 0));
    return 0;
}

extern "C" int sw2_batch_set_launch(sw2_batch *b, int blocks_eta, int blocks_uv, int steps_per_graph) {
    if (blocks_eta < 0 || blocks_uv < 0 || steps_per_graph < 1 || steps_per_graph > 256) {
        sw2_set_error("launch shape out of range");
        return -1;
    }
    if ((blocks_eta || blocks_uv) && b->shape != 0) {
        sw2_set_error("persistent workgroups run %d x %d tiles only", FTX, FTY);
        return -1;
    }
    HIPCHECK(hipStreamSynchronize(b->st));
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        (void)hipGraphDestroy(b->graph);
        b->gexec = NULL;
        b->graph = NULL;
    }
    b->blocks_eta = blocks_eta;
    b->blocks_uv = blocks_uv;
    b->gsteps = steps_per_graph;
    return 0;
}

extern "C" int sw2_batch_get_launch(sw2_batch *b, int *blocks_eta, int *blocks_uv, int *steps_per_graph) {
    if (blocks_eta) *blocks_eta = b->blocks_eta;
    if (blocks_uv) *blocks_uv = b->blocks_uv;
    if (steps_per_graph) *steps_per_graph = b->gsteps;
    return 0;
}

extern "C" int sw2_batch_set_tiles(sw2_batch *b, int tile_x, int tile_y) {
    int k = -1;
    for (int q = 0; q < NBATCH_SHAPES; q++)
        if (BATCH_SHAPES[q].tx == tile_x && BATCH_SHAPES[q].ty == tile_y) k = q;
    if (k < 0) {
        sw2_set_error("no batch kernels for %d x %d tiles (32 x 16, 64 x 8 and 64 x 16 are built)", tile_x, tile_y);
        return -1;
    }
    if (k != 0 && (b->blocks_eta || b->blocks_uv)) {
        sw2_set_error("persistent workgroups run %d x %d tiles only", FTX, FTY);
        return -1;
    }
    if (k == b->shape) return 0;
    HIPCHECK(hipStreamSynchronize(b->st));
    if (b->gexec) {  // the captured step launches the other shape's kernels
        (void)hipGraphExecDestroy(b->gexec);
        (void)hipGraphDestroy(b->graph);
        b->gexec = NULL;
        b->graph = NULL;
    }
    b->shape = k;
    for (int m = 0; m < b->cap; m++) {
        if (!b->sess[m]) continue;
        b->h_mem[m].ntx = (b->sess[m]->g.nx + tile_x - 1) / tile_x;
        b->h_mem[m].nty = (b->sess[m]->g.ny + tile_y - 1) / tile_y;
    }
    batch_retile(b);
    return 0;
}

extern "C" int sw2_batch_get_tiles(sw2_batch *b, int *tile_x, int *tile_y) {
    if (tile_x) *tile_x = BATCH_SHAPES[b->shape].tx;
    if (tile_y) *tile_y = BATCH_SHAPES[b->shape].ty;
    return 0;
}

// after a directly launched batch step: the storms of every member that
// took it, into the buffer holding its new height, and its tracer and
// height checks after them (as a single mixed session orders them).  The member's session counts its
// steps and time as a single session does (time by repeated addition of
// Dt), so its storms are those of its own run.
static int batch_storms(sw2_batch *b) {
    for (int m = 0; m < b->cap; m++) {
        sw2_session *s = b->sess[m];
        const BMember &M = b->h_mem[m];
        if (!s || !M.late || b->paused[m] || s->n >= M.nlast) continue;
        const int n = s->n + 1;
        const double t = s->t + s->Dt;
        MField eta_new = M.eta[((n - M.nstart - 1) & 1) ^ 1];
        if (s->perturbs_dirty) {
            k_zero_pert<<<(s->S + 255) / 256, 256, 0, b->st>>>(s->fl, n, s->S, s->pert);
            s->perturbs_dirty = 0;
        }
        int eta_changed = 0;
        if (host_perturbations_on(s, n, t, &eta_changed, b->st, eta_new)) return -1;
        if (s->tracer_active)
            k_check_mixed<<<cells_grid(s->g), cells_block(), 0, b->st>>>(s->g, s->fl, s->m_tr, nullptr, n, 2);
        k_check_mixed<<<cells_grid(s->g), cells_block(), 0, b->st>>>(s->g, s->fl, eta_new, s->d_mt + MT_HBAR * s->g.R, n, 3);
        s->n = n;
        s->t = t;
    }
    return 0;
}

// the host half of a replayed window of G steps with storm members: each
// storm member's storms at each step it will take in the window, evaluated
// on the host as a single mixed session evaluates them step by step (SW2's
// expressions and libm), gathered into the window table and one block of
// cell lists, and copied to the device ahead of the graph that reads them.
// The members' step counts, times and mark states advance here, as
// batch_storms advances them step by step; a member that fails a guard in
// the window is held back on the device by its guard record (the kernels'
// gates), as it is in the direct launches.
static int batch_storm_window(sw2_batch *b, int G) {
    const int nl = b->graph_nl;
    size_t used = 0;
    auto take = [&](size_t bytes) -> long {  // 8-aligned room in the host assembly, as an offset
        const size_t off = (used + 7) & ~(size_t)7;
        if (off + bytes > b->build_cap) {
            const size_t cap = (off + bytes) * 2 + (1 << 16);
            char *nb = (char *)realloc(b->h_build, cap);
            if (!nb) return -1;
            b->h_build = nb;
            b->build_cap = cap;
        }
        used = off + bytes;
        return (long)off;
    };
    for (int q = 0; q < G; q++) {
        int e = 0;
        for (int m = 0; m < b->cap; m++) {
            sw2_session *s = b->sess[m];
            const BMember &M = b->h_mem[m];
            if (!s || !M.late) continue;
            SEntry &E = b->h_wt[(size_t)q * nl + e++];
            memset(&E, 0, sizeof E);
            if (b->paused[m] || s->n >= M.nlast) continue;
            const int n = s->n + 1;
            const double t = s->t + s->Dt;
            E.g = s->g;
            E.fl = s->fl;
            E.n = n;
            E.S = s->S;
            E.pert = s->pert;
            E.eta = M.eta[((n - M.nstart - 1) & 1) ^ 1];
            E.tr = s->m_tr;
            E.tr_on = s->tracer_active;
            E.hbar = s->d_mt + MT_HBAR * s->g.R;
            E.zero = s->perturbs_dirty;
            s->perturbs_dirty = 0;
            const int K = sw2_gaussians_sparse(s->SW, s->Dt, t, SMAXK, s->inj);
            if (K < 0) return -1;
            for (int k = 0; k < K; k++) {
                const sw2_inj &x = s->inj[k];
                if (x.n == 0) continue;
                const long oi = take(sizeof(int) * (size_t)x.n), oc = oi < 0 ? -1 : take(sizeof(double) * (size_t)x.n),
                           ot = oc < 0 ? -1 : take(sizeof(double) * (size_t)x.n);
                if (ot < 0) {
                    sw2_set_error("out of host memory for the storm cells of a window");
                    return -1;
                }
                memcpy(b->h_build + oi, x.idx, sizeof(int) * (size_t)x.n);
                memcpy(b->h_build + oc, x.inc, sizeof(double) * (size_t)x.n);
                memcpy(b->h_build + ot, x.trv, sizeof(double) * (size_t)x.n);
                const int c = E.nk++;
                E.cnt[c] = x.n;
                E.has_inc[c] = x.has_inc;
                // offsets into the cell block until it has its device address
                E.idx[c] = (const int *)(uintptr_t)oi;
                E.inc[c] = (const double *)(uintptr_t)oc;
                E.trv[c] = (const double *)(uintptr_t)ot;
                s->perturbs_dirty = 1;
            }
            s->n = n;
            s->t = t;
        }
    }
    if (used > b->cells_cap) {
        HIPCHECK(hipStreamSynchronize(b->st));  // the last replay may still read the old block
        if (b->d_cells) HIPCHECK(hipFree(b->d_cells));
        b->d_cells = NULL;
        b->cells_cap = 0;
        const size_t cap = used + used / 2 + (1 << 16);
        HIPCHECK(hipMalloc((void **)&b->d_cells, cap));
        b->cells_cap = cap;
    }
    const size_t nent = (size_t)G * nl;
    for (size_t k = 0; k < nent; k++) {
        SEntry &E = b->h_wt[k];
        for (int c = 0; c < E.nk; c++) {
            E.idx[c] = (const int *)(b->d_cells + (uintptr_t)E.idx[c]);
            E.inc[c] = (const double *)(b->d_cells + (uintptr_t)E.inc[c]);
            E.trv[c] = (const double *)(b->d_cells + (uintptr_t)E.trv[c]);
        }
    }
    const int si = b->stage_i;
    b->stage_i ^= 1;
    const size_t tab = sizeof(SEntry) * nent, need = tab + used;
    if (b->stage_ev[si]) HIPCHECK(hipEventSynchronize(b->stage_ev[si]));
    else HIPCHECK(hipEventCreateWithFlags(&b->stage_ev[si], hipEventDisableTiming));
    if (need > b->stage_cap[si]) {
        if (b->h_stage[si]) HIPCHECK(hipHostFree(b->h_stage[si]));
        b->h_stage[si] = NULL;
        b->stage_cap[si] = 0;
        const size_t cap = need + need / 2 + (1 << 16);
        HIPCHECK(hipHostMalloc((void **)&b->h_stage[si], cap, hipHostMallocDefault));
        b->stage_cap[si] = cap;
    }
    memcpy(b->h_stage[si], b->h_wt, tab);
    if (used) memcpy(b->h_stage[si] + tab, b->h_build, used);
    HIPCHECK(hipMemcpyAsync(b->d_wt, b->h_stage[si], tab, hipMemcpyHostToDevice, b->st));
    if (used) HIPCHECK(hipMemcpyAsync(b->d_cells, b->h_stage[si] + tab, used, hipMemcpyHostToDevice, b->st));
    HIPCHECK(hipEventRecord(b->stage_ev[si], b->st));
    return 0;
}

extern "C" int sw2_batch_step(sw2_batch *b, int nsteps) {
    if (b->nmem == 0 || nsteps <= 0) return 0;
    if (batch_upload(b)) return -1;
    // whole graphs of gsteps steps (members with storms included: the host
    // fills each replay's storm window ahead of it), the remainder launched
    // directly, storms after each step
    if (nsteps >= b->gsteps && batch_capture(b)) return -1;
    int q = 0;
    for (; q + b->gsteps <= nsteps; q += b->gsteps) {
        if (b->graph_nl && batch_storm_window(b, b->gsteps)) return -1;
        HIPCHECK(hipGraphLaunch(b->gexec, b->st));
    }
    for (; q < nsteps; q++) {
        batch_launch_step(b);
        if (b->storms && batch_storms(b)) return -1;
    }
    HIPCHECK(hipGetLastError());
    return 0;
}

extern "C" int sw2_batch_sync(sw2_batch *b) {
    HIPCHECK(hipStreamSynchronize(b->st));
    return 0;
}

extern "C" int sw2_batch_member(sw2_batch *b, int id, int *steps_done, int *done, int *failed_step) {
    if (id < 0 || id >= b->cap || !b->sess[id]) {
        sw2_set_error("no member %d", id);
        return -1;
    }
    BState S;
    HIPCHECK(hipMemcpyAsync(&S, b->d_st + id, sizeof S, hipMemcpyDeviceToHost, b->st));
    HIPCHECK(hipStreamSynchronize(b->st));
    if (steps_done) *steps_done = S.nstep;
    if (done) *done = S.done;
    if (failed_step) *failed_step = sw2_session_failed_step(b->sess[id], NULL, 0);
    return 0;
}

// hold a member (its tiles taken out of the launches, its step count kept)
// or release it; between steps only: the caller has synchronised the batch
extern "C" int sw2_batch_pause(sw2_batch *b, int id, int paused) {
    if (id < 0 || id >= b->cap || !b->sess[id]) {
        sw2_set_error("no member %d", id);
        return -1;
    }
    const int v = paused ? 1 : 0;
    HIPCHECK(hipMemcpyAsync(reinterpret_cast<char *>(b->d_st + id) + offsetof(BState, paused), &v, sizeof v,
                            hipMemcpyHostToDevice, b->st));
    HIPCHECK(hipStreamSynchronize(b->st));
    if (b->paused[id] != v) {
        b->paused[id] = v;
        batch_retile(b);
    }
    return 0;
}

// ids of the members whose runs ended (last step reached or a guard failed)
extern "C" int sw2_batch_finished(sw2_batch *b, int *ids, int maxids) {
    BState *h = (BState *)malloc(sizeof(BState) * b->cap);
    if (hipMemcpyAsync(h, b->d_st, sizeof(BState) * b->cap, hipMemcpyDeviceToHost, b->st) != hipSuccess ||
        hipStreamSynchronize(b->st) != hipSuccess) {
        free(h);
        sw2_set_error("reading member states failed");
        return -1;
    }
    int n = 0;
    for (int m = 0; m < b->cap; m++)
        if (b->sess[m] && h[m].done && n < maxids) ids[n++] = m;
    free(h);
    return n;
}

extern "C" int sw2_batch_download(sw2_batch *b, int id, sw *SW) {
    int steps = 0;
    if (sw2_batch_member(b, id, &steps, NULL, NULL)) return -1;
    sw2_session *s = b->sess[id];
    const BMember &M = b->h_mem[id];
    const int p = (steps - M.nstart) & 1;  // buffer holding the state after `steps` steps
    s->m_eta = M.eta[p];
    s->m_eta2 = M.eta[p ^ 1];
    s->m_u = M.u[p];
    s->m_u2 = M.u[p ^ 1];
    s->m_v = M.v[p];
    s->m_v2 = M.v[p ^ 1];
    if (!M.late) { /* a storm member's session already counts its steps and time */
        s->n = steps;
        s->t = SW->t0 + steps * s->Dt;
    }
    return sw2_session_download(s, SW);
}

extern "C" int sw2_batch_energy(sw2_batch *b, int id, double *ke, double *ape, double *mass) {
    int steps = 0;
    if (sw2_batch_member(b, id, &steps, NULL, NULL)) return -1;
    sw2_session *s = b->sess[id];
    const BMember &M = b->h_mem[id];
    const int p = (steps - M.nstart) & 1;
    s->m_eta = M.eta[p];
    s->m_u = M.u[p];
    s->m_v = M.v[p];
    return sw2_session_energy(s, ke, ape, mass);
}

extern "C" int sw2_batch_remove(sw2_batch *b, int id) {
    if (id < 0 || id >= b->cap || !b->sess[id]) {
        sw2_set_error("no member %d", id);
        return -1;
    }
    HIPCHECK(hipStreamSynchronize(b->st));
    BState z{0, 1, 0};
    HIPCHECK(hipMemcpy(b->d_st + id, &z, sizeof z, hipMemcpyHostToDevice));
    if (b->h_mem[id].late) b->storms--;
    if (b->h_mem[id].tr_on) b->tracers--;
    memset(&b->h_mem[id], 0, sizeof(BMember));
    batch_retile(b);
    // the member's device arrays are released only after the table that
    // still names them has been replaced
    if (batch_upload(b)) return -1;
    sw2_session_destroy(b->sess[id]);
    b->sess[id] = NULL;
    b->used -= b->foot[id];
    b->foot[id] = 0;
    b->nmem--;
    return 0;
}

extern "C" int sw2_batch_count(sw2_batch *b, size_t *used_bytes) {
    if (used_bytes) *used_bytes = b->used;
    return b->nmem;
}

extern "C" void sw2_batch_destroy(sw2_batch *b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (int m = 0; m < b->cap; m++)
        if (b->sess && b->sess[m]) sw2_session_destroy(b->sess[m]);
    if (b->d_mem) (void)hipFree(b->d_mem);
    if (b->d_st) (void)hipFree(b->d_st);
    if (b->d_flags) (void)hipFree(b->d_flags);
    if (b->d_tmap) (void)hipFree(b->d_tmap);
    free(b->h_tmap);
    if (b->d_wt) (void)hipFree(b->d_wt);
    if (b->d_cells) (void)hipFree(b->d_cells);
    free(b->h_wt);
    free(b->h_build);
    for (int k = 0; k < 2; k++) {
        if (b->h_stage[k]) (void)hipHostFree(b->h_stage[k]);
        if (b->stage_ev[k]) (void)hipEventDestroy(b->stage_ev[k]);
    }
    if (b->st) (void)hipStreamDestroy(b->st);
    free(b->sess);
    free(b->foot);
    free(b->h_mem);
    free(b->paused);
    free(b);
}

#include "sw2_fine.inc"
