// Element kernels for gfx950: T10/T4 residual and residual + analytic tangent.
//
// Arithmetic follows the reference element and law routines:
//   F per quadrature point, local fibre frame   CBElementSolidT10.cpp:266-288
//   Guccione PK2                                CBConstitutiveModelGuccione.cpp:103-138
//   Holzapfel PK2 (Jm = pow(J, -2/3) = 1)       CBConstitutiveModelHolzapfel.cpp:118-154
//   active stress T/sqrt(I4f) f(x)f             CBTensionModel.h:37-50
//   residual f_a = detJ/24 sum_q dN_a . P_q     CBElementSolidT10.cpp:620-693 (T4: :147-210)
// The tangent is the exact derivative of that residual (the reference uses central
// finite differences instead, CBElementSolidT10.cpp:511-567):
//   K(a,b)_ij = w sum_q sum_{m,P} G_a,m M^{ij}_{mP} G_b,P,  G = B^-1 dN/dX,
//   M^{ij}_{mP} = (B B^T)_ij S_Pm + sum_{I,K} Ft_iI Ft_jK D_{Im,KP},  Ft = B F_loc,
// with D = dS/dE of the law in the fibre frame, evaluated in closed form.
//
// Lane mapping of the tangent kernel: 64-lane workgroup, 7 elements, 9 lanes per
// element, lane (i, j) owns K(a,b)_ij for all node pairs (a, b) and computes only
// the 9 entries M^{ij} it needs, so no LDS and no barriers are used.
#include <climits>

#include <type_traits>

#include "cmg_internal.h"

namespace cmg {
namespace {

// ---------------------------------------------------------------------------
// analytic time functions (CBAnalyticFunction.h:53-145), steady Linear window
// ---------------------------------------------------------------------------
__device__ double dh_raw(const double* fn, double t) {
    double c = pow(t / fn[6], fn[8]);
    c = c / (1.0 + c);
    double r = pow(t / fn[7], fn[9]);
    r = 1.0 / (1.0 + r);
    return c * r;
}

__device__ double analytic(int type, const double* fn, double time) {
#pragma clang fp contract(off)
    const double start = fn[0], stop = fn[1];
    switch (type) {
    case CMG_FN_LINEAR:
        if (time < start) return 0.0;
        if (time < stop) return (time - start) * 1.0 / (stop - start);
        return fn[2] != 0.0 ? 1.0 : 0.0;
    case CMG_FN_SINUS:
        if ((time >= start) && (time <= stop)) return 0.5 - 0.5 * cos(fn[3] * (time - start) + fn[4]);
        return 0.0;
    case CMG_FN_SINUSDRIVER:
        if ((time >= start) && (time <= stop)) return sin(3.14159 * (time - start) / (stop - start));
        return 0.0;
    case CMG_FN_DOUBLEHILL: {
        if ((time >= start) && (time <= stop)) {
            double t = time - fn[10];
            t = (t > 0.0) ? fmod(t, fn[5]) : 0.0;
            return dh_raw(fn, t) / fn[11];
        }
        return 0.0;
    }
    default:
        return 0.0;
    }
}

// Time part of the active tension of an element.  FromFunction: Tmax g(t - t_act)
// (CBTensionFromFunction.cpp:59-61), the whole tension.  TanH
// (CBTensionModelTanH.cpp:71-97): t is wrapped into [0, cycleLength] by repeated
// subtraction, t_s = clamp(t - t_act - t_emd, 0, t_dur) is returned in ts and the
// value is Tmax S_peak tanh^2((t_dur - t_s)/tau_r); the stretch-dependent factor
// phi tanh^2(t_s/tau_c(phi)) is applied per quadrature point by eval_law.
// fn of TanH: cycleLength, l0, t_emd, S_peak, t_dur, tau_c0, ld, ld_up, tau_r, kff,
// startTime (added to the activation time).
__device__ __forceinline__ double tension_of(const DevMaterial& m, double time, double tact, double& ts) {
    ts = 0.0;
    if (m.tension_fn == CMG_FN_NONE) return 0.0;
    if (m.tension_fn == CMG_FN_LAND17) return 1.0;   // evaluated per quadrature point from the element state
    if (m.tension_fn == CMG_FN_TANH) {
        const double* f = m.fn;
        double t = time;
        while (t > f[0]) t -= f[0];
        ts = fmax(fmin(t - (tact + f[10]) - f[2], f[4]), 0.0);
        const double tr = tanh((f[4] - ts) / f[8]);
        return m.tension_max * f[3] * (tr * tr);
    }
    return m.tension_max * analytic(m.tension_fn, m.fn, time - tact);
}

// ---------------------------------------------------------------------------
// Land17 (CBTensionModelLand17.cpp): one explicit Euler step of the cross-bridge ODEs from
// the element's committed state per evaluation (CalcActiveTension, :298-372), the tension
// Tmax * max(0, Ta + F_d) and its derivative in the fibre stretch.
// ---------------------------------------------------------------------------
struct L17 {
    const double* p;   // parameter block [40] of the element's material
    double st[8];      // committed XS XW TRPN TmBlocked ZetaS ZetaW Cd lambda
    double dt_ms;      // (time - commit time) in ms
    double cai;        // Cai(time) of the element
    double tmax;
};

// CaCoppini / CaBrixius (:55-245): 1 ms samples since activation (+ startTime), wrapped
// into the cycle, index clipped to [0, 1000]
__device__ __forceinline__ double land17_cai(const double* p, const double* tab, double time, double tact) {
    double t = time - (tact + p[22]);
    while (t > p[23]) t = t - p[23];
    int index = (int)(t / 1e-3);
    if (index < 0) index = 0;
    index = min(1000, index);
    return tab[(p[24] != 0.0 ? 1001 : 0) + index];
}

__device__ __forceinline__ void load_l17(const ElemArgs& a, int64_t e, int mi, const DevMaterial& m, L17& s) {
    s.p = a.l17_params + kLand17Params * mi;
#pragma unroll
    for (int k = 0; k < 8; k++) s.st[k] = a.l17_state[(int64_t)k * a.E + e];
    s.dt_ms = a.l17_dt_ms;
    s.cai = land17_cai(s.p, a.l17_cai, a.time, a.tact[e]);
    s.tmax = m.tension_max;
}

// tension at stretch lam; dT/dlambda into *dT and the advanced state into next[8] when
// not null (constants of the constructor, :262-272)
__device__ inline double land17_tension(const L17& s, double lam, double* dT, double* next) {
#pragma clang fp contract(off)
    const double* p = s.p;
    const double TRPNn = p[0], TRPNk = p[1], Ca50_ = p[2], ku = p[3], nTM = p[4], TRPN50 = p[5], kuw = p[6];
    const double rs = p[8], rw = p[9], gammas = p[10], gammaw = p[11], phi = p[12], Aeff = p[13];
    const double beta0 = p[14], beta1 = p[15], Tref = p[16], a = p[17], k = p[18], eta_l = p[19], eta_s = p[20];
    const double kws = p[21] * p[7];
    const double cw = phi * kuw * ((1.0 - rs) * (1.0 - rw)) / ((1.0 - rs) * rw);
    const double cs = phi * kws * ((1.0 - rs) * rw) / rs;
    const double kwu = kuw * (1.0 / rw - 1.0) - kws;
    const double ksu = (kws * rw * (1.0 / rs - 1.0));
    const double A = Aeff * rs / ((1.0 - rs) * rw + rs);
    const double dt = s.dt_ms;
    const double XS0 = s.st[0], XW0 = s.st[1], TRPN0 = s.st[2], TB0 = s.st[3], ZS0 = s.st[4], ZW0 = s.st[5];
    const double Cd0 = s.st[6], lam0 = s.st[7];
    const bool rate = !(dt == 0.0 || p[25] == 0.0);
    const double dldt = rate ? (lam - lam0) / dt : 0.0;
    const double ddldt = rate ? 1.0 / dt : 0.0;
    const double lam_m = (lam < 1.2) ? lam : 1.2;
    const double dlm = (lam < 1.2) ? 1.0 : 0.0;
    const double lm87 = (lam_m < 0.87) ? lam_m : 0.87;
    const double dlm87 = (lam_m < 0.87) ? dlm : 0.0;
    const double hraw = 1.0 + beta0 * (lam_m + lm87 - 1.87);
    const double h = (hraw < 0.0) ? 0.0 : hraw;
    const double dh = (hraw < 0.0) ? 0.0 : beta0 * (dlm + dlm87);
    const double XU = (1.0 - TB0) - XW0 - XS0;
    const double xb_ws = kws * XW0, xb_uw = kuw * XU, xb_wu = kwu * XW0, xb_su = ksu * XS0;
    const double zs_pos = (ZS0 > 0.0) ? ZS0 : 0.0;
    const double zs_neg = (ZS0 < -1.0) ? (-ZS0 - 1.0) : 0.0;
    const double zs = fmax(zs_pos, zs_neg);
    const double xb_su_gamma = gammas * zs * XS0;
    const double XS = XS0 + dt * (xb_ws - xb_su - xb_su_gamma);
    const double gr_w = (ZW0 < 0.0) ? -ZW0 : ZW0;
    const double xb_wu_gamma = gammaw * gr_w * XW0;
    const double XW = XW0 + dt * (xb_uw - xb_wu - xb_ws - xb_wu_gamma);
    const double ZS = ZS0 + dt * (A * dldt - cs * ZS0);
    const double ZW = ZW0 + dt * (A * dldt - cw * ZW0);
    const double Taraw = h * (Tref / rs) * ((ZS + 1.0) * XS + ZW * XW);
    const double Ta = (Taraw < 0.0) ? 0.0 : Taraw;
    const double C_s = (lam - 1.0) - Cd0;
    const double F_d = a * k * C_s;
    const double Traw = Ta + F_d;
    const double Tension = (Traw < 0.0) ? 0.0 : Traw;
    if (dT) {
        const double dZ = dt * (A * ddldt);
        const double dTa = (Taraw < 0.0) ? 0.0 : (Tref / rs) * (dh * ((ZS + 1.0) * XS + ZW * XW) + h * (dZ * XS + dZ * XW));
        *dT = (Traw < 0.0) ? 0.0 : s.tmax * (dTa + a * k);
    }
    if (next) {
        const double Ca50 = Ca50_ + beta1 * (lam_m - 1.0);
        const double TRPN = TRPN0 + dt * (TRPNk * (pow(s.cai / Ca50, TRPNn) * (1.0 - TRPN0) - TRPN0));
        const double ktm_block = ku * pow(TRPN50, nTM) * 0.5 / (0.5 - rs * 0.5 - (1.0 - rs) * rw * 0.5);
        const double TRPN_NP = fmin(100.0, pow(TRPN, (-nTM / 2.0)));
        const double eta = (C_s > 0.0) ? eta_l : eta_s;
        next[0] = XS;
        next[1] = XW;
        next[2] = TRPN;
        next[3] = TB0 + dt * (ktm_block * TRPN_NP * XU - ku * pow(TRPN, nTM / 2.0) * TB0);
        next[4] = ZS;
        next[5] = ZW;
        next[6] = Cd0 + dt * (k * C_s / eta);
        next[7] = lam;
    }
    return s.tmax * Tension;
}

// ---------------------------------------------------------------------------
// 3x3 helpers, row-major, reference term order (Matrix3.h:290-303,482-540)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mul3(const double* a, const double* b, double* r) {
#pragma clang fp contract(off)
    r[0] = a[0] * b[0] + a[1] * b[3] + a[2] * b[6];
    r[1] = a[0] * b[1] + a[1] * b[4] + a[2] * b[7];
    r[2] = a[0] * b[2] + a[1] * b[5] + a[2] * b[8];
    r[3] = a[3] * b[0] + a[4] * b[3] + a[5] * b[6];
    r[4] = a[3] * b[1] + a[4] * b[4] + a[5] * b[7];
    r[5] = a[3] * b[2] + a[4] * b[5] + a[5] * b[8];
    r[6] = a[6] * b[0] + a[7] * b[3] + a[8] * b[6];
    r[7] = a[6] * b[1] + a[7] * b[4] + a[8] * b[7];
    r[8] = a[6] * b[2] + a[7] * b[5] + a[8] * b[8];
}

__device__ __forceinline__ void tr3(const double* a, double* r) {
    r[0] = a[0]; r[1] = a[3]; r[2] = a[6];
    r[3] = a[1]; r[4] = a[4]; r[5] = a[7];
    r[6] = a[2]; r[7] = a[5]; r[8] = a[8];
}

__device__ __forceinline__ double det3(const double* e) {
#pragma clang fp contract(off)
    return (e[0] * e[4] * e[8]) + (e[3] * e[7] * e[2]) + (e[6] * e[1] * e[5]) - (e[2] * e[4] * e[6]) -
           (e[5] * e[7] * e[0]) - (e[8] * e[3] * e[1]);
}

__device__ __forceinline__ void inv3(const double* e, double* r) {
#pragma clang fp contract(off)
    double d = det3(e);
    if (d == 0.0) {
        for (int i = 0; i < 9; i++) r[i] = INFINITY;
        return;
    }
    r[0] = (-e[5] * e[7] + e[4] * e[8]) / d;
    r[1] = (e[2] * e[7] - e[1] * e[8]) / d;
    r[2] = (-e[2] * e[4] + e[1] * e[5]) / d;
    r[3] = (e[5] * e[6] - e[3] * e[8]) / d;
    r[4] = (-e[2] * e[6] + e[0] * e[8]) / d;
    r[5] = (e[2] * e[3] - e[0] * e[5]) / d;
    r[6] = (-e[4] * e[6] + e[3] * e[7]) / d;
    r[7] = (e[1] * e[6] - e[0] * e[7]) / d;
    r[8] = (-e[1] * e[3] + e[0] * e[4]) / d;
}

// ---------------------------------------------------------------------------
// constitutive laws in the fibre frame
// ---------------------------------------------------------------------------
// Everything the tangent needs besides S, for
//   M_mP = BBt S_Pm + czz (Zu)_m (Zw)_P + codot [(u.Zw) Z_mP + (Zu)_P (Zw)_m]
//        + Guccione: ceq/2 (d_mP h_m + u_P w_m b_Pm) + 2 ceq (Yu)_m (Yw)_P
//        + Holzapfel: (Xu)_m w_P + mfs a_m b_P
//        + d00 u_0 w_0 [m=P=0] + d11 u_1 w_1 [m=P=1]
struct Law {
    double S[9];
    double Z[9];     // C^-1
    double czz, codot;
    double ceq;      // Guccione C e^Q
    double Y[9];     // Guccione b (.) E
    double X[9];     // Holzapfel 2gb(I - I1/3 Z) - 2g/3 Z
    double d00, d11, mfs;
};

// returns false for det F <= 0 (CBStatus::CORRUPT_ELEMENT)
__device__ __forceinline__ bool eval_law(const DevMaterial& m, const double* F, double Ta, double ts,
                                         bool want_tangent, Law& L, const L17* l17 = nullptr) {
#pragma clang fp contract(off)
    const double J = det3(F);
    if (!m.ignore_corrupt && (J <= 0.0)) return false;
    double Ft[9], C[9];
    tr3(F, Ft);
    mul3(Ft, F, C);
    inv3(C, L.Z);
    L.d00 = 0.0;
    L.d11 = 0.0;
    L.mfs = 0.0;
    if (m.law == CMG_LAW_GUCCIONE) {
        const double Cg = m.p[0], bf = m.p[1], bt = m.p[2], bfs = m.p[3], k = m.p[4];
        double e[9];
        for (int i = 0; i < 9; i++) e[i] = 0.5 * (C[i] - ((i % 4 == 0) ? 1.0 : 0.0));
        const double Q = bf * e[0] * e[0] + bt * (e[4] * e[4] + e[8] * e[8] + e[5] * e[5] + e[7] * e[7]) +
                         bfs * (e[1] * e[1] + e[3] * e[3] + e[2] * e[2] + e[6] * e[6]);
        const double expQ = exp(Q);
        double* p = L.S;
        p[0] = Cg * bf * e[0] * expQ;
        p[1] = Cg * bfs * e[1] * expQ;
        p[2] = Cg * bfs * e[2] * expQ;
        p[3] = Cg * bfs * e[3] * expQ;
        p[4] = Cg * bt * e[4] * expQ;
        p[5] = Cg * bt * e[5] * expQ;
        p[6] = Cg * bfs * e[6] * expQ;
        p[7] = Cg * bt * e[7] * expQ;
        p[8] = Cg * bt * e[8] * expQ;
        const double vol = k * (J - 1.0) * J;
        for (int i = 0; i < 9; i++) p[i] += vol * L.Z[i];
        if (want_tangent) {
            const double b[9] = {bf, bfs, bfs, bfs, bt, bt, bfs, bt, bt};
            L.ceq = Cg * expQ;
            for (int i = 0; i < 9; i++) L.Y[i] = b[i] * e[i];
            L.czz = k * (2.0 * J - 1.0) * J;
            L.codot = -k * (J * J - J);
        }
    } else {
        const double a = m.p[0], b = m.p[1], af = m.p[2], bf = m.p[3], as = m.p[4], bs = m.p[5];
        const double afs = m.p[6], bfs = m.p[7], kappa = m.p[8], kh = m.p[9];
        const double I1 = C[0] + C[4] + C[8];
        const double I4f = C[0], I4s = C[4];
        const double I8 = 0.5 * (C[1] + C[3]);
        double Hf, Hs, dHf, dHs;
        if (kh == 0.0) {
            Hf = (I4f >= 1.0) ? 1.0 : 0.0;
            Hs = (I4s >= 1.0) ? 1.0 : 0.0;
            dHf = 0.0;
            dHs = 0.0;
        } else {
            Hf = 1.0 / (1.0 + exp(-kh * (I4f - 1.0)));
            Hs = 1.0 / (1.0 + exp(-kh * (I4s - 1.0)));
            dHf = kh * Hf * (1.0 - Hf);
            dHs = kh * Hs * (1.0 - Hs);
        }
        const double cvol = (kappa / 2.0) * (J - 1.0 / J) * J;
        const double g = 1.0 * a * exp(b * (1.0 * I1 - 3.0));
        const double ef = exp(bf * (I4f - 1.0) * (I4f - 1.0));
        const double es = exp(bs * (I4s - 1.0) * (I4s - 1.0));
        const double efs = exp(bfs * I8 * I8);
        const double cf = 2.0 * af * Hf * (I4f - 1.0) * ef;
        const double cs = 2.0 * as * Hs * (I4s - 1.0) * es;
        const double cfs = afs * I8 * efs;
        for (int i = 0; i < 9; i++) {
            const double id = (i % 4 == 0) ? 1.0 : 0.0;
            L.S[i] = cvol * L.Z[i] + g * (id - 1.0 / 3.0 * I1 * L.Z[i]);
        }
        L.S[0] += cf;
        L.S[4] += cs;
        L.S[1] += cfs;
        L.S[3] += cfs;
        if (want_tangent) {
            L.czz = kappa * J * J;
            L.codot = -(kappa / 2.0) * (J * J - 1.0) + g * I1 / 3.0;
            for (int i = 0; i < 9; i++) {
                const double id = (i % 4 == 0) ? 1.0 : 0.0;
                L.X[i] = 2.0 * g * b * (id - I1 / 3.0 * L.Z[i]) - (2.0 * g / 3.0) * L.Z[i];
            }
            L.d00 = 2.0 * 2.0 * af * ef * (dHf * (I4f - 1.0) + Hf * (1.0 + 2.0 * bf * (I4f - 1.0) * (I4f - 1.0)));
            L.d11 = 2.0 * 2.0 * as * es * (dHs * (I4s - 1.0) + Hs * (1.0 + 2.0 * bs * (I4s - 1.0) * (I4s - 1.0)));
            L.mfs = afs * efs * (1.0 + 2.0 * bfs * I8 * I8);
        }
    }
    // active stress with the default stress coefficients (only ff): T / sqrt(I4f)
    if (Ta != 0.0) {
        const double I4f = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
        const double rs = 1. / sqrt(I4f);
        if (m.tension_fn == CMG_FN_LAND17) {
            // S_ff = kff T(lambda) / lambda with the element's cross-bridge state
            if (l17) {
                const double lam = sqrt(I4f);
                double dT = 0.0;
                const double T = land17_tension(*l17, lam, want_tangent ? &dT : nullptr, nullptr);
                const double kff = l17->p[26];
                L.S[0] += kff * T * rs;
                if (want_tangent) L.d00 += kff * (dT * rs * rs - T * rs * rs * rs);
            }
        } else if (m.tension_fn == CMG_FN_TANH) {
            // T(lambda) = Ta phi tanh^2(t_s / tau_c), phi = tanh(ld (lambda - l0)),
            // tau_c = tau_c0 + ld_up (1 - phi), clipped at 0; S_ff = kff T / lambda
            const double* f = m.fn;
            const double lam = sqrt(I4f);
            const double phi = tanh(f[6] * (lam - f[1]));
            const double tauc = f[5] + f[7] * (1.0 - phi);
            const double th = tanh(ts / tauc);
            const double T = Ta * phi * (th * th);
            if (T > 0.0) {
                L.S[0] += f[9] * T * rs;
                if (want_tangent) {
                    const double dphi = f[6] * (1.0 - phi * phi);
                    const double dth = (1.0 - th * th) * (ts / (tauc * tauc)) * (f[7] * dphi);
                    const double dT = Ta * (dphi * (th * th) + phi * 2.0 * th * dth);   // dT / dlambda
                    L.d00 += f[9] * (dT * rs * rs - T * rs * rs * rs);
                }
            }
        } else {
            L.S[0] += Ta * rs;
            if (want_tangent) L.d00 += -Ta * rs / I4f;
        }
    }
    return true;
}

// M^{ij}_{mP} for the lane's spatial pair (i, j)
template <typename TT>
__device__ __forceinline__ void lane_M(const DevMaterial& m, const Law& L, const double* Ftil, const double* B,
                                       int ii, int jj, TT* M) {
    const double* u = Ftil + 3 * ii;
    const double* w = Ftil + 3 * jj;
    const double bbt = B[3 * ii] * B[3 * jj] + B[3 * ii + 1] * B[3 * jj + 1] + B[3 * ii + 2] * B[3 * jj + 2];
    double Zu[3], Zw[3];
    for (int r = 0; r < 3; r++) {
        Zu[r] = L.Z[3 * r] * u[0] + L.Z[3 * r + 1] * u[1] + L.Z[3 * r + 2] * u[2];
        Zw[r] = L.Z[3 * r] * w[0] + L.Z[3 * r + 1] * w[1] + L.Z[3 * r + 2] * w[2];
    }
    const double uZw = u[0] * Zw[0] + u[1] * Zw[1] + u[2] * Zw[2];
    double Mm[9];
    for (int mm = 0; mm < 3; mm++)
        for (int P = 0; P < 3; P++)
            Mm[3 * mm + P] = bbt * L.S[3 * P + mm] + L.czz * Zu[mm] * Zw[P] +
                             L.codot * (uZw * L.Z[3 * mm + P] + Zu[P] * Zw[mm]);
    if (m.law == CMG_LAW_GUCCIONE) {
        const double bf = m.p[1], bt = m.p[2], bfs = m.p[3];
        const double b[9] = {bf, bfs, bfs, bfs, bt, bt, bfs, bt, bt};
        double Yu[3], Yw[3], h[3];
        for (int r = 0; r < 3; r++) {
            Yu[r] = L.Y[3 * r] * u[0] + L.Y[3 * r + 1] * u[1] + L.Y[3 * r + 2] * u[2];
            Yw[r] = L.Y[3 * r] * w[0] + L.Y[3 * r + 1] * w[1] + L.Y[3 * r + 2] * w[2];
            h[r] = u[0] * w[0] * b[r] + u[1] * w[1] * b[3 + r] + u[2] * w[2] * b[6 + r];
        }
        const double hc = 0.5 * L.ceq, oc = 2.0 * L.ceq;
        for (int mm = 0; mm < 3; mm++)
            for (int P = 0; P < 3; P++)
                Mm[3 * mm + P] += hc * (u[P] * w[mm] * b[3 * P + mm] + ((mm == P) ? h[mm] : 0.0)) +
                                  oc * Yu[mm] * Yw[P];
    } else {
        double Xu[3];
        for (int r = 0; r < 3; r++) Xu[r] = L.X[3 * r] * u[0] + L.X[3 * r + 1] * u[1] + L.X[3 * r + 2] * u[2];
        const double av[3] = {u[1], u[0], 0.0}, bv[3] = {w[1], w[0], 0.0};
        for (int mm = 0; mm < 3; mm++)
            for (int P = 0; P < 3; P++) Mm[3 * mm + P] += Xu[mm] * w[P] + L.mfs * av[mm] * bv[P];
        Mm[4] += L.d11 * u[1] * w[1];
    }
    Mm[0] += L.d00 * u[0] * w[0];
    for (int r = 0; r < 9; r++) M[r] = (TT)Mm[r];
}

template <int NPE>
__device__ __forceinline__ void load_g(const ElemArgs& a, int64_t e, int q, double (*g)[3]) {
    const int64_t E = a.E;
#pragma unroll
    for (int k = 0; k < NPE; k++)
#pragma unroll
        for (int d = 0; d < 3; d++) g[k][d] = a.dNdX[((int64_t)(q * NPE + k) * 3 + d) * E + e];
}

template <int NPE>
__device__ __forceinline__ void local_F(const double* x, const double (*g)[3], const double* B, double* Binv,
                                        double* Floc) {
#pragma clang fp contract(off)
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NPE; k++) s += g[k][j] * x[3 * k + i];
            F[3 * i + j] = s;
        }
    double Bt[9], BinvT[9], T1[9];
    inv3(B, Binv);
    tr3(B, Bt);
    tr3(Binv, BinvT);
    mul3(Bt, F, T1);
    mul3(T1, BinvT, Floc);
}

// ---------------------------------------------------------------------------
// parameter-set batch: blockIdx.y selects the set (see ElemArgs)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool batch_select(const ElemArgs& a0, ElemArgs& a) {
    const int64_t set = blockIdx.y;
    if (a0.active && !a0.active[set]) return false;
    a = a0;
    a.x += set * a0.set_x;
    a.fe += set * a0.set_fe;
    a.Ke = static_cast<char*>(a0.Ke) + set * a0.set_Ke_bytes;
    a.mats += set * a0.set_mats;
    a.corrupt += set;
    return true;
}

__device__ __forceinline__ bool batch_select(const SurfArgs& a0, SurfArgs& a) {
    const int64_t set = blockIdx.y;
    if (a0.active && !a0.active[set]) return false;
    a = a0;
    a.x += set * a0.set_x;
    a.fs += set * a0.set_fs;
    if (a.Ks) a.Ks += set * a0.set_Ks;
    if (a.vol_partial) a.vol_partial += set * a0.set_vol;
    return true;
}

// ---------------------------------------------------------------------------
// Land17 state commit of an accepted step: the element's state advances from the
// evaluation at its last quadrature point (the reference's last CalcActiveTension call
// of the step, SaveStateVariablesAsNeeded, CBTensionModelLand17.cpp:274-293)
// ---------------------------------------------------------------------------
template <int NPE, int NQP>
__global__ __launch_bounds__(128) void k_land17_commit(const ElemArgs a, double* state) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    const int mi = a.matidx ? a.matidx[e] : 0;
    const DevMaterial& m = a.mats[mi];
    if (m.tension_fn != CMG_FN_LAND17) return;
    const int64_t E = a.E;
    double x[3 * NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
        const int64_t n = a.conn[(int64_t)k * E + e];
        x[3 * k] = a.x[3 * n];
        x[3 * k + 1] = a.x[3 * n + 1];
        x[3 * k + 2] = a.x[3 * n + 2];
    }
    constexpr int q = NQP - 1;
    double g[NPE][3];
    load_g<NPE>(a, e, q, g);
    double B[9], Binv[9], Floc[9];
#pragma unroll
    for (int r = 0; r < 9; r++) B[r] = a.bases[(int64_t)(q * 9 + r) * E + e];
    local_F<NPE>(x, g, B, Binv, Floc);
    L17 s;
    load_l17(a, e, mi, m, s);
    double next[8];
    land17_tension(s, sqrt(Floc[0] * Floc[0] + Floc[3] * Floc[3] + Floc[6] * Floc[6]), nullptr, next);
#pragma unroll
    for (int k = 0; k < 8; k++) state[(int64_t)k * E + e] = next[k];
}

// ---------------------------------------------------------------------------
// residual only: one thread per element, FP64
// ---------------------------------------------------------------------------
// LAND17: instantiation for contexts with a Land17 material (its state and
// parameters are read only there; the other instantiations carry none of it)
template <int NPE, int NQP, bool LAND17 = false>
__global__ __launch_bounds__(kResidualBlock) void k_residual(const ElemArgs a0) {
#pragma clang fp contract(off)
    ElemArgs a;
    if (!batch_select(a0, a)) return;
    const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int64_t e = (int64_t)bid * kResidualBlock + threadIdx.x;
    if (e >= a.E) return;
    const int64_t E = a.E;
    const DevMaterial& m = a.mats[a.matidx ? a.matidx[e] : 0];
    double ts;
    const double Ta = tension_of(m, a.time, a.tact[e], ts);
    L17 l17s;
    const L17* l17 = nullptr;
    if (LAND17 && m.tension_fn == CMG_FN_LAND17) {
        load_l17(a, e, a.matidx ? a.matidx[e] : 0, m, l17s);
        l17 = &l17s;
    }
    double x[3 * NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
        const int64_t n = a.conn[(int64_t)k * E + e];
        x[3 * k] = a.x[3 * n];
        x[3 * k + 1] = a.x[3 * n + 1];
        x[3 * k + 2] = a.x[3 * n + 2];
    }
    double t[NPE][3];
#pragma unroll
    for (int k = 0; k < NPE; k++) t[k][0] = t[k][1] = t[k][2] = 0.0;
    for (int q = 0; q < NQP; q++) {
        double g[NPE][3];
        load_g<NPE>(a, e, q, g);
        double B[9], Binv[9], Floc[9];
#pragma unroll
        for (int r = 0; r < 9; r++) B[r] = a.bases[(int64_t)(q * 9 + r) * E + e];
        local_F<NPE>(x, g, B, Binv, Floc);
        Law L;
        if (!eval_law(m, Floc, Ta, ts, false, L, l17)) {
            atomicMin(a.corrupt, (int)e);
            return;
        }
        // stress = (F S)^T (T10, :649) or S F^T (T4, :176); S is symmetric so both agree
        double FS[9], P1[9], P[9], BinvT[9], Bt[9], St[9];
        mul3(Floc, L.S, FS);
        tr3(FS, St);
        tr3(Binv, BinvT);
        tr3(B, Bt);
        mul3(BinvT, St, P1);
        mul3(P1, Bt, P);
#pragma unroll
        for (int k = 0; k < NPE; k++)
#pragma unroll
            for (int c = 0; c < 3; c++) t[k][c] += g[k][0] * P[0 + c] + g[k][1] * P[3 + c] + g[k][2] * P[6 + c];
    }
    const double V = a.w[e];
    double* fe = a.fe + e * (3 * NPE);
#pragma unroll
    for (int k = 0; k < NPE; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) fe[3 * k + c] = V * t[k][c];
}

// ---------------------------------------------------------------------------
// residual + tangent: 9 lanes per element (see file header)
// ---------------------------------------------------------------------------
template <int NPE, int NQP, typename TT, bool LAND17 = false>
__global__ __launch_bounds__(64) void k_residual_tangent(const ElemArgs a0) {
    ElemArgs a;
    if (!batch_select(a0, a)) return;
    const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    const int slot = lane / 9;
    const int ij = lane - 9 * slot;
    const int ii = ij / 3, jj = ij - 3 * (ij / 3);
    const int64_t e = (int64_t)bid * kTanElemsPerWave + slot;
    if (slot >= kTanElemsPerWave || e >= a.E) return;
    const int64_t E = a.E;
    const DevMaterial& m = a.mats[a.matidx ? a.matidx[e] : 0];
    double ts;
    const double Ta = tension_of(m, a.time, a.tact[e], ts);
    L17 l17s;
    const L17* l17 = nullptr;
    if (LAND17 && m.tension_fn == CMG_FN_LAND17) {
        load_l17(a, e, a.matidx ? a.matidx[e] : 0, m, l17s);
        l17 = &l17s;
    }
    double x[3 * NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
        const int64_t n = a.conn[(int64_t)k * E + e];
        x[3 * k] = a.x[3 * n];
        x[3 * k + 1] = a.x[3 * n + 1];
        x[3 * k + 2] = a.x[3 * n + 2];
    }
    TT acc[NPE * NPE];
#pragma unroll
    for (int r = 0; r < NPE * NPE; r++) acc[r] = (TT)0;
    double fr[NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) fr[k] = 0.0;
    for (int q = 0; q < NQP; q++) {
        double B[9], Binv[9], Floc[9];
        TT G[NPE][3];
        Law L;
        {
            double g[NPE][3];
            load_g<NPE>(a, e, q, g);
#pragma unroll
            for (int r = 0; r < 9; r++) B[r] = a.bases[(int64_t)(q * 9 + r) * E + e];
            local_F<NPE>(x, g, B, Binv, Floc);
            if (!eval_law(m, Floc, Ta, ts, true, L, l17)) {
                atomicMin(a.corrupt, (int)e);
                return;
            }
            // G_k = B^-1 dN_k/dX
#pragma unroll
            for (int k = 0; k < NPE; k++)
#pragma unroll
                for (int r = 0; r < 3; r++)
                    G[k][r] = (TT)(Binv[3 * r] * g[k][0] + Binv[3 * r + 1] * g[k][1] + Binv[3 * r + 2] * g[k][2]);
            if (jj == 0) {
                // residual component ii in the reference form (FP64, exact dN/dX)
#pragma clang fp contract(off)
                double FS[9], St[9], BinvT[9], Bt[9], P1[9], P[9];
                mul3(Floc, L.S, FS);
                tr3(FS, St);
                tr3(Binv, BinvT);
                tr3(B, Bt);
                mul3(BinvT, St, P1);
                mul3(P1, Bt, P);
#pragma unroll
                for (int k = 0; k < NPE; k++) fr[k] += g[k][0] * P[ii] + g[k][1] * P[3 + ii] + g[k][2] * P[6 + ii];
            }
        }
        double Ftil[9];
        mul3(B, Floc, Ftil);
        TT M[9];
        lane_M<TT>(m, L, Ftil, B, ii, jj, M);
#pragma unroll
        for (int b = 0; b < NPE; b++) {
            const TT mg0 = M[0] * G[b][0] + M[1] * G[b][1] + M[2] * G[b][2];
            const TT mg1 = M[3] * G[b][0] + M[4] * G[b][1] + M[5] * G[b][2];
            const TT mg2 = M[6] * G[b][0] + M[7] * G[b][1] + M[8] * G[b][2];
#pragma unroll
            for (int aa = 0; aa < NPE; aa++) {
                // three chained fused multiply-adds into the accumulator
                TT t = acc[aa * NPE + b];
                t = fma(G[aa][0], mg0, t);
                t = fma(G[aa][1], mg1, t);
                acc[aa * NPE + b] = fma(G[aa][2], mg2, t);
            }
        }
    }
    const double V = a.w[e];
    TT* Ke = static_cast<TT*>(a.Ke) + e * (9 * NPE * NPE);
    const TT Vt = (TT)V;
#pragma unroll
    for (int r = 0; r < NPE * NPE; r++) Ke[9 * r + ij] = Vt * acc[r];
    if (jj == 0) {
        double* fe = a.fe + e * (3 * NPE);
#pragma unroll
        for (int k = 0; k < NPE; k++) fe[3 * k + ii] = V * fr[k];
    }
}

// ---------------------------------------------------------------------------
// residual + tangent, shared quadrature-point stage (v2)
//
// Stage A: 28 lanes (7 elements x 4 quadrature points) each evaluate one point
// -- F, the law, the reference-order residual partial, G = B^-1 dN/dX and the
// data the tangent needs -- and park it in LDS.  Stage B: 63 lanes (7 elements x
// 9 spatial pairs (i, j)) build M^{ij} from LDS and contract it with G into
// K(a,b)_ij.  NSPLIT > 1 splits the node range a into NSPLIT passes so fewer
// accumulators stay live (more passes, less register pressure).  BSPLIT > 1
// gives each (element, i, j) BSPLIT lanes, one per part of the node range b, in a
// workgroup of BSPLIT waves: fewer accumulators per lane, more waves per SIMD.
// ---------------------------------------------------------------------------
constexpr int kQRec = 51;   // Ft(9) BBt(9) S(9) Z(9) Y|X(9) czz codot ceq d00 d11 mfs

// STAGED: the accumulators of the workgroup's elements go through LDS (the quadrature
// records' space, free after the contraction) in rounds of RA node rows a, and the
// workgroup writes each element's rows as one contiguous run of 16-byte stores --
// instead of every lane storing its 3x3 entries 72 bytes at a time, 7200 bytes apart.
template <int NPE, int NQP, typename TT, int NSPLIT, int BSPLIT = 1, int GROUPS = 1, bool LAND17 = false,
          bool STAGED = false>
__global__ __launch_bounds__(64 * BSPLIT * GROUPS) void k_residual_tangent_v2(const ElemArgs a0) {
    constexpr int EPW = kTanElemsPerWave;
    constexpr int EPG = EPW * GROUPS;     // elements per workgroup
    static_assert(EPG * NQP <= 64, "stage A runs in the first wave");
    static_assert(!STAGED || NSPLIT == 1, "the staged store keeps every accumulator of the lane");
    ElemArgs a;
    if (!batch_select(a0, a)) return;   // uniform per workgroup: no barrier below is skipped by part of it
    constexpr int ROW = 9 * NPE;                                        // K_e values of one node row a
    constexpr int RA = STAGED ? (sizeof(TT) == 8 ? NPE / 2 : NPE) : 1;   // node rows per staging round
    constexpr size_t kRecBytes = sizeof(double) * EPG * NQP * 3 * NPE + sizeof(TT) * EPG * NQP * (NPE * 3 + kQRec);
    constexpr size_t kStageBytes = STAGED ? sizeof(TT) * EPG * RA * ROW : 0;
    __shared__ __align__(16) unsigned char smem[kRecBytes > kStageBytes ? kRecBytes : kStageBytes];
    auto& sT = *reinterpret_cast<double (*)[EPG][NQP][3 * NPE]>(smem);
    auto& sG = *reinterpret_cast<TT (*)[EPG][NQP][NPE * 3]>(smem + sizeof(double) * EPG * NQP * 3 * NPE);
    auto& sQ = *reinterpret_cast<TT (*)[EPG][NQP][kQRec]>(smem + sizeof(double) * EPG * NQP * 3 * NPE +
                                                          sizeof(TT) * EPG * NQP * NPE * 3);
    const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    const int64_t E = a.E;
    const int64_t e0 = (int64_t)bid * EPG;
    // ---- stage A ----
    if (lane < EPG * NQP) {
        // element index fastest: the SoA loads (dN/dX, bases, connectivity) of consecutive
        // lanes hit consecutive elements of one row
        const int q = lane / EPG, slot = lane - EPG * (lane / EPG);
        const int64_t e = e0 + slot;
        if (e < E) {
            const DevMaterial& m = a.mats[a.matidx ? a.matidx[e] : 0];
            double ts;
            const double Ta = tension_of(m, a.time, a.tact[e], ts);
            L17 l17s;
            const L17* l17 = nullptr;
            if (LAND17 && m.tension_fn == CMG_FN_LAND17) {
                load_l17(a, e, a.matidx ? a.matidx[e] : 0, m, l17s);
                l17 = &l17s;
            }
            double x[3 * NPE];
#pragma unroll
            for (int k = 0; k < NPE; k++) {
                const int64_t n = a.conn[(int64_t)k * E + e];
                x[3 * k] = a.x[3 * n];
                x[3 * k + 1] = a.x[3 * n + 1];
                x[3 * k + 2] = a.x[3 * n + 2];
            }
            double g[NPE][3];
            load_g<NPE>(a, e, q, g);
            double B[9], Binv[9], Floc[9];
#pragma unroll
            for (int r = 0; r < 9; r++) B[r] = a.bases[(int64_t)(q * 9 + r) * E + e];
            local_F<NPE>(x, g, B, Binv, Floc);
            Law L;
            if (!eval_law(m, Floc, Ta, ts, true, L, l17)) atomicMin(a.corrupt, (int)e);
            {
#pragma clang fp contract(off)
                double FS[9], St[9], BinvT[9], Bt[9], P1[9], P[9];
                mul3(Floc, L.S, FS);
                tr3(FS, St);
                tr3(Binv, BinvT);
                tr3(B, Bt);
                mul3(BinvT, St, P1);
                mul3(P1, Bt, P);
#pragma unroll
                for (int k = 0; k < NPE; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        sT[slot][q][3 * k + c] = g[k][0] * P[0 + c] + g[k][1] * P[3 + c] + g[k][2] * P[6 + c];
            }
#pragma unroll
            for (int k = 0; k < NPE; k++)
#pragma unroll
                for (int r = 0; r < 3; r++)
                    sG[slot][q][3 * k + r] =
                        (TT)(Binv[3 * r] * g[k][0] + Binv[3 * r + 1] * g[k][1] + Binv[3 * r + 2] * g[k][2]);
            double Ft[9];
            mul3(B, Floc, Ft);
            TT* rec = sQ[slot][q];
#pragma unroll
            for (int r = 0; r < 9; r++) {
                rec[r] = (TT)Ft[r];
                rec[9 + r] = (TT)(B[3 * (r / 3)] * B[3 * (r % 3)] + B[3 * (r / 3) + 1] * B[3 * (r % 3) + 1] +
                                  B[3 * (r / 3) + 2] * B[3 * (r % 3) + 2]);
                rec[18 + r] = (TT)L.S[r];
                rec[27 + r] = (TT)L.Z[r];
                rec[36 + r] = (TT)((m.law == CMG_LAW_GUCCIONE) ? L.Y[r] : L.X[r]);
            }
            rec[45] = (TT)L.czz;
            rec[46] = (TT)L.codot;
            rec[47] = (TT)((m.law == CMG_LAW_GUCCIONE) ? L.ceq : 0.0);
            rec[48] = (TT)L.d00;
            rec[49] = (TT)L.d11;
            rec[50] = (TT)L.mfs;
        }
    }
    __syncthreads();
    // ---- stage B ----
    // wave w serves element group w / BSPLIT and part w % BSPLIT of the b range
    const int bl = lane % 64, bh = (lane / 64) % BSPLIT, grp = lane / (64 * BSPLIT);
    const int lslot = bl / 9;
    const int ij = bl - 9 * lslot;
    const int ii = ij / 3, jj = ij - 3 * (ij / 3);
    const int slot = grp * EPW + lslot;
    const int64_t e = e0 + slot;
    const bool valid = lslot < EPW && e < E;
    if (!STAGED && !valid) return;
    constexpr int AH = NPE / NSPLIT;
    constexpr int BH = NPE / BSPLIT;
    const int b0 = bh * BH;
    TT keep[STAGED ? AH * BH : 1];   // V K_e entries of this lane for the staged store
    if (valid) {
        const DevMaterial& m = a.mats[a.matidx ? a.matidx[e] : 0];
        const bool guc = m.law == CMG_LAW_GUCCIONE;
        TT bmat[9];
        {
            const TT bf = (TT)m.p[1], bt = (TT)m.p[2], bfs = (TT)m.p[3];
            const TT bb[9] = {bf, bfs, bfs, bfs, bt, bt, bfs, bt, bt};
#pragma unroll
            for (int r = 0; r < 9; r++) bmat[r] = bb[r];
        }
        const double V = a.w[e];
        TT* Ke = static_cast<TT*>(a.Ke) + e * (9 * NPE * NPE);
#pragma unroll 1
        for (int half = 0; half < NSPLIT; half++) {
            TT acc[AH * BH];
#pragma unroll
            for (int r = 0; r < AH * BH; r++) acc[r] = (TT)0;
#pragma unroll 1
            for (int q = 0; q < NQP; q++) {
                const TT* rec = sQ[slot][q];
                TT u[3], w[3], Z[9], S[9], YX[9];
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    u[r] = rec[3 * ii + r];
                    w[r] = rec[3 * jj + r];
                }
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    S[r] = rec[18 + r];
                    Z[r] = rec[27 + r];
                    YX[r] = rec[36 + r];
                }
                const TT bbt = rec[9 + ij], czz = rec[45], codot = rec[46], ceq = rec[47];
                const TT d00 = rec[48], d11 = rec[49], mfs = rec[50];
                TT Zu[3], Zw[3], Xu[3], Yw[3];
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    Zu[r] = Z[3 * r] * u[0] + Z[3 * r + 1] * u[1] + Z[3 * r + 2] * u[2];
                    Zw[r] = Z[3 * r] * w[0] + Z[3 * r + 1] * w[1] + Z[3 * r + 2] * w[2];
                    Xu[r] = YX[3 * r] * u[0] + YX[3 * r + 1] * u[1] + YX[3 * r + 2] * u[2];
                    Yw[r] = YX[3 * r] * w[0] + YX[3 * r + 1] * w[1] + YX[3 * r + 2] * w[2];
                }
                const TT uZw = u[0] * Zw[0] + u[1] * Zw[1] + u[2] * Zw[2];
                TT M[9];
#pragma unroll
                for (int mm = 0; mm < 3; mm++)
#pragma unroll
                    for (int P = 0; P < 3; P++)
                        M[3 * mm + P] = bbt * S[3 * P + mm] + czz * Zu[mm] * Zw[P] +
                                        codot * (uZw * Z[3 * mm + P] + Zu[P] * Zw[mm]);
                if (guc) {
                    const TT hc = (TT)0.5 * ceq, oc = (TT)2.0 * ceq;
                    TT h[3];
#pragma unroll
                    for (int r = 0; r < 3; r++)
                        h[r] = u[0] * w[0] * bmat[r] + u[1] * w[1] * bmat[3 + r] + u[2] * w[2] * bmat[6 + r];
#pragma unroll
                    for (int mm = 0; mm < 3; mm++)
#pragma unroll
                        for (int P = 0; P < 3; P++)
                            M[3 * mm + P] += hc * (u[P] * w[mm] * bmat[3 * P + mm] + ((mm == P) ? h[mm] : (TT)0)) +
                                             oc * Xu[mm] * Yw[P];
                } else {
                    const TT av[3] = {u[1], u[0], (TT)0}, bv[3] = {w[1], w[0], (TT)0};
#pragma unroll
                    for (int mm = 0; mm < 3; mm++)
#pragma unroll
                        for (int P = 0; P < 3; P++) M[3 * mm + P] += Xu[mm] * w[P] + mfs * av[mm] * bv[P];
                    M[4] += d11 * u[1] * w[1];
                }
                M[0] += d00 * u[0] * w[0];
                TT G[NPE][3];
#pragma unroll
                for (int k = 0; k < NPE; k++)
#pragma unroll
                    for (int r = 0; r < 3; r++) G[k][r] = sG[slot][q][3 * k + r];
                // b range of this lane with a compile-time offset (no dynamic register indexing)
                auto contract = [&](auto B0) {
                    constexpr int kB0 = decltype(B0)::value;
#pragma unroll
                    for (int bb = 0; bb < BH; bb++) {
                        const int b = kB0 + bb;
                        const TT mg0 = M[0] * G[b][0] + M[1] * G[b][1] + M[2] * G[b][2];
                        const TT mg1 = M[3] * G[b][0] + M[4] * G[b][1] + M[5] * G[b][2];
                        const TT mg2 = M[6] * G[b][0] + M[7] * G[b][1] + M[8] * G[b][2];
#pragma unroll
                        for (int aa = 0; aa < AH; aa++) {
                            const int ga = half * AH + aa;
                            // three chained fused multiply-adds into the accumulator
                            TT t = acc[aa * BH + bb];
                            t = fma(G[ga][0], mg0, t);
                            t = fma(G[ga][1], mg1, t);
                            acc[aa * BH + bb] = fma(G[ga][2], mg2, t);
                        }
                    }
                };
                if (BSPLIT == 1 || bh == 0)
                    contract(std::integral_constant<int, 0>());
                else
                    contract(std::integral_constant<int, BH>());
            }
            const TT Vt = (TT)V;
            if constexpr (STAGED) {
#pragma unroll
                for (int r = 0; r < AH * BH; r++) keep[r] = Vt * acc[r];
            } else {
#pragma unroll
                for (int aa = 0; aa < AH; aa++)
#pragma unroll
                    for (int bb = 0; bb < BH; bb++)
                        Ke[9 * ((half * AH + aa) * NPE + b0 + bb) + ij] = Vt * acc[aa * BH + bb];
            }
        }
        if (jj == 0 && bh == 0) {
            double* fe = a.fe + e * (3 * NPE);
#pragma unroll
            for (int k = 0; k < NPE; k++) {
                double t = 0.0;
#pragma unroll
                for (int q = 0; q < NQP; q++) t += sT[slot][q][3 * k + ii];
                fe[3 * k + ii] = V * t;
            }
        }
    }   // valid
    if constexpr (STAGED) {
        // 16-byte units: RA rows of one element are RA * ROW values, a whole number of units,
        // and every element's and round's start is 16-byte aligned (ROW * sizeof(TT) * RA and
        // 9 * NPE * NPE * sizeof(TT) are multiples of 16 for the shapes instantiated)
        constexpr int VPU = 16 / sizeof(TT);
        constexpr int UNITS = RA * ROW / VPU;
        static_assert((RA * ROW) % VPU == 0 && (RA * ROW * sizeof(TT)) % 16 == 0 &&
                          (9 * NPE * NPE * sizeof(TT)) % 16 == 0, "16-byte runs");
        constexpr int NT = 64 * BSPLIT * GROUPS;
        auto& stage = *reinterpret_cast<TT (*)[EPG][RA][ROW]>(smem);
        const int nel = (int)((E - e0 < EPG) ? (E - e0) : EPG);
#pragma unroll
        for (int r0 = 0; r0 < NPE; r0 += RA) {
            __syncthreads();   // the records (first round) or the previous round are consumed
            if (valid) {
#pragma unroll
                for (int aa = 0; aa < RA; aa++)
#pragma unroll
                    for (int bb = 0; bb < BH; bb++) stage[slot][aa][9 * (b0 + bb) + ij] = keep[(r0 + aa) * BH + bb];
            }
            __syncthreads();
            for (int u = lane; u < nel * UNITS; u += NT) {
                const int sl = u / UNITS, off = u - sl * UNITS;
                const uint4 v = reinterpret_cast<const uint4*>(&stage[sl][0][0])[off];
                TT* dst = static_cast<TT*>(a.Ke) + (e0 + sl) * (9 * NPE * NPE) + r0 * ROW;
                reinterpret_cast<uint4*>(dst)[off] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// surface pressure loads (CBElementSurfaceT3.cpp:257-315, CBElementSurfaceT6.cpp:248-303)
// with exact tangents by forward-mode dual numbers, and the volume contribution
// (CBElementSurfaceT3.cpp:151-191, CBElementSurfaceT6.cpp:191-241)
// ---------------------------------------------------------------------------
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(double s, Dual b) { return {s * b.v, s * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
    double s = sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(Dual a) { return a.v; }
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }

template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__device__ void t3_forces(const T* x, double p, T* f) {
    T ab[3], ac[3], n[3];
    for (int i = 0; i < 3; i++) {
        ab[i] = x[i] - x[3 + i];
        ac[i] = x[i] - x[6 + i];
    }
    cross3(ab, ac, n);
    T nn = dsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    T area = 0.5 * nn;
    if (val(nn) != 0.0)
        for (int i = 0; i < 3; i++) n[i] = n[i] / nn;
    T s = (-p * (1. / 3.)) * area;
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) f[3 * a + c] = s * n[c];
}

template <typename T>
__device__ void tri_normal(const T* v0, const T* v1, const T* v2, T* n, T& area) {
    T a[3], b[3];
    for (int i = 0; i < 3; i++) {
        a[i] = v1[i] - v0[i];
        b[i] = v2[i] - v0[i];
    }
    cross3(a, b, n);
    T nn = dsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    area = 0.5 * nn;
    if (val(nn) != 0.0)
        for (int i = 0; i < 3; i++) n[i] = n[i] / nn;
}

template <typename T>
__device__ void t6_forces(const T* x, double p, T* f) {
    T t1[3], t2[3], t3[3], t4[3], A1, A2, A3, A4;
    tri_normal(x + 0, x + 9, x + 15, t1, A1);
    tri_normal(x + 9, x + 12, x + 15, t2, A2);
    tri_normal(x + 9, x + 3, x + 12, t3, A3);
    tri_normal(x + 15, x + 12, x + 6, t4, A4);
    T area = A1 + A2 + A3 + A4;
    T nv[3][3];
    for (int i = 0; i < 3; i++) {
        nv[0][i] = t1[i] + t2[i];
        nv[1][i] = t3[i] + t2[i];
        nv[2][i] = t4[i] + t2[i];
    }
    for (int k = 0; k < 3; k++) {
        T nn = dsqrt(nv[k][0] * nv[k][0] + nv[k][1] * nv[k][1] + nv[k][2] * nv[k][2]);
        if (val(nn) != 0.0)
            for (int i = 0; i < 3; i++) nv[k][i] = nv[k][i] / nn;
    }
    const double l[3][3] = {{2. / 3., 1. / 6., 1. / 6.}, {1. / 6., 2. / 3., 1. / 6.}, {1. / 6., 1. / 6., 2. / 3.}};
    T s = (-p / 3.0) * area;
    for (int i = 0; i < 6; i++) {
        double N[3];
        for (int k = 0; k < 3; k++) {
            const double l1 = l[k][0], l2 = l[k][1], l3 = l[k][2];
            N[k] = (i == 0) ? l1 * (2 * l1 - 1)
                 : (i == 1) ? l2 * (2 * l2 - 1)
                 : (i == 2) ? l3 * (2 * l3 - 1)
                 : (i == 3) ? 4 * l1 * l2
                 : (i == 4) ? 4 * l1 * l3
                            : 4 * l2 * l3;
        }
        for (int c = 0; c < 3; c++) f[3 * i + c] = s * (N[0] * nv[0][c] + N[1] * nv[1][c] + N[2] * nv[2][c]);
    }
}

template <int NPS>
__global__ __launch_bounds__(64) void k_surface(const SurfArgs a0, bool tangent) {
    SurfArgs a;
    if (!batch_select(a0, a)) return;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.S) return;
    const int32_t* cn = a.conn + s * NPS;
    const double p = a.pressure[a.group[s]];
    double x[3 * NPS];
    uint8_t bc[3 * NPS];
    for (int k = 0; k < NPS; k++) {
        const int64_t n = cn[k];
        const uint8_t mask = a.bcmask ? a.bcmask[n] : 0;
        for (int c = 0; c < 3; c++) {
            x[3 * k + c] = a.x[3 * n + c];
            bc[3 * k + c] = (mask >> c) & 1;
        }
    }
    double f[3 * NPS];
    if (NPS == 3)
        t3_forces<double>(x, p, f);
    else
        t6_forces<double>(x, p, f);
    for (int r = 0; r < 3 * NPS; r++) a.fs[s * (3 * NPS) + r] = bc[r] ? 0.0 : f[r];
    if (a.vol_partial) {
        double v;
        if (NPS == 3) {
            double aa[3] = {x[3] - x[0], x[4] - x[1], x[5] - x[2]};
            double bb[3] = {x[6] - x[0], x[7] - x[1], x[8] - x[2]};
            double c[3];
            cross3(aa, bb, c);
            v = (x[0] * c[0] + x[1] * c[1] + x[2] * c[2]) / 6.0;
        } else {
            double a1[3], a2[3], a3[3], a4[3], a5[3];
            for (int i = 0; i < 3; i++) {
                a1[i] = x[9 + i] - x[i];
                a2[i] = x[3 + i] - x[i];
                a3[i] = x[12 + i] - x[i];
                a4[i] = x[6 + i] - x[i];
                a5[i] = x[15 + i] - x[i];
            }
            double c1[3], c2[3], c3[3], c4[3];
            cross3(a1, a2, c1);
            cross3(a2, a3, c2);
            cross3(a3, a4, c3);
            cross3(a4, a5, c4);
            double v1 = x[0] * c1[0] + x[1] * c1[1] + x[2] * c1[2];
            double v2 = x[0] * c2[0] + x[1] * c2[1] + x[2] * c2[2];
            double v3 = x[0] * c3[0] + x[1] * c3[1] + x[2] * c3[2];
            double v4 = x[0] * c4[0] + x[1] * c4[1] + x[2] * c4[2];
            v = (v1 + v2 + v3 + v4) / 6.0;
        }
        a.vol_partial[s] = v;
    }
    if (tangent && a.Ks) {
        constexpr int n = 3 * NPS;
        double* K = a.Ks + s * (n * n);
        for (int c = 0; c < n; c++) {
            Dual xd[n], fd[n];
            for (int i = 0; i < n; i++) xd[i] = {x[i], (i == c) ? 1.0 : 0.0};
            if (NPS == 3)
                t3_forces<Dual>(xd, p, fd);
            else
                t6_forces<Dual>(xd, p, fd);
            for (int r = 0; r < n; r++) K[n * r + c] = (bc[r] || bc[c]) ? 0.0 : fd[r].d;
        }
    }
}

// dN/dX at the 4 quadrature points (SoA) and detJ/24 per element,
// CBElementSolidT10::CalcShapeFunctionDerivatives (:94-264)
__device__ double t10_shape_point(const double* X, double l1, double l2, double l3, double l4, double* out) {
#pragma clang fp contract(off)
    double Jx[4], Jy[4], Jz[4];
    const double* c[10];
    for (int k = 0; k < 10; k++) c[k] = X + 3 * k;
    for (int d = 0; d < 3; d++) {
        double* J = (d == 0) ? Jx : (d == 1) ? Jy : Jz;
        J[0] = 4.0 * (c[0][d] * (l1 - 0.25) + c[4][d] * l2 + c[6][d] * l3 + c[7][d] * l4);
        J[1] = 4.0 * (c[4][d] * l1 + c[1][d] * (l2 - 0.25) + c[5][d] * l3 + c[8][d] * l4);
        J[2] = 4.0 * (c[6][d] * l1 + c[5][d] * l2 + c[2][d] * (l3 - 0.25) + c[9][d] * l4);
        J[3] = 4.0 * (c[7][d] * l1 + c[8][d] * l2 + c[9][d] * l3 + c[3][d] * (l4 - 0.25));
    }
#define D_(v, i, j) (v[(i)-1] - v[(j)-1])
    const double detJ = D_(Jx, 2, 1) * (D_(Jy, 2, 3) * D_(Jz, 3, 4) - D_(Jy, 3, 4) * D_(Jz, 2, 3)) +
                        D_(Jx, 3, 2) * (D_(Jy, 3, 4) * D_(Jz, 1, 2) - D_(Jy, 1, 2) * D_(Jz, 3, 4)) +
                        D_(Jx, 4, 3) * (D_(Jy, 1, 2) * D_(Jz, 2, 3) - D_(Jy, 2, 3) * D_(Jz, 1, 2));
    const double av[4] = {D_(Jy, 4, 2) * D_(Jz, 3, 2) - D_(Jy, 3, 2) * D_(Jz, 4, 2),
                          D_(Jy, 3, 1) * D_(Jz, 4, 3) - D_(Jy, 3, 4) * D_(Jz, 1, 3),
                          D_(Jy, 2, 4) * D_(Jz, 1, 4) - D_(Jy, 1, 4) * D_(Jz, 2, 4),
                          D_(Jy, 1, 3) * D_(Jz, 2, 1) - D_(Jy, 1, 2) * D_(Jz, 3, 1)};
    const double bv[4] = {D_(Jx, 3, 2) * D_(Jz, 4, 2) - D_(Jx, 4, 2) * D_(Jz, 3, 2),
                          D_(Jx, 4, 3) * D_(Jz, 3, 1) - D_(Jx, 1, 3) * D_(Jz, 3, 4),
                          D_(Jx, 1, 4) * D_(Jz, 2, 4) - D_(Jx, 2, 4) * D_(Jz, 1, 4),
                          D_(Jx, 2, 1) * D_(Jz, 1, 3) - D_(Jx, 3, 1) * D_(Jz, 1, 2)};
    const double cv[4] = {D_(Jx, 4, 2) * D_(Jy, 3, 2) - D_(Jx, 3, 2) * D_(Jy, 4, 2),
                          D_(Jx, 3, 1) * D_(Jy, 4, 3) - D_(Jx, 3, 4) * D_(Jy, 1, 3),
                          D_(Jx, 2, 4) * D_(Jy, 1, 4) - D_(Jx, 1, 4) * D_(Jy, 2, 4),
                          D_(Jx, 1, 3) * D_(Jy, 2, 1) - D_(Jx, 1, 2) * D_(Jy, 3, 1)};
#undef D_
    const double fac = 4.0 / detJ;
    const double* comp[3] = {av, bv, cv};
    for (int d = 0; d < 3; d++) {
        const double* v = comp[d];
        out[0 + d] = fac * (l1 - 0.25) * v[0];
        out[3 + d] = fac * (l2 - 0.25) * v[1];
        out[6 + d] = fac * (l3 - 0.25) * v[2];
        out[9 + d] = fac * (l4 - 0.25) * v[3];
        out[12 + d] = fac * (l1 * v[1] + l2 * v[0]);
        out[15 + d] = fac * (l2 * v[2] + l3 * v[1]);
        out[18 + d] = fac * (l3 * v[0] + l1 * v[2]);
        out[21 + d] = fac * (l1 * v[3] + l4 * v[0]);
        out[24 + d] = fac * (l2 * v[3] + l4 * v[1]);
        out[27 + d] = fac * (l3 * v[3] + l4 * v[2]);
    }
    return detJ;
}

__global__ void k_t10_shape(int64_t E, const int32_t* conn, const double* Xg, double* dNdX, double* w) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double X[30], out[30];
    for (int k = 0; k < 10; k++) {
        const int64_t n = conn[e * 10 + k];
        for (int c = 0; c < 3; c++) X[3 * k + c] = Xg[3 * n + c];
    }
    const double qa = (5 + 3 * sqrt(5.0)) / 20, qb = (5 - sqrt(5.0)) / 20;
    const double L[5][4] = {{qa, qb, qb, qb}, {qb, qa, qb, qb}, {qb, qb, qa, qb}, {qb, qb, qb, qa},
                            {0.25, 0.25, 0.25, 0.25}};
    double detJ = 0.0;
    for (int q = 0; q < 5; q++) {
        detJ = t10_shape_point(X, L[q][0], L[q][1], L[q][2], L[q][3], out);
        if (q < 4)
            for (int r = 0; r < 30; r++) dNdX[(int64_t)(q * 30 + r) * E + e] = out[r];
    }
    w[e] = detJ / 24.0;
}

// Per-element state for export, one thread per element: the passive PK2 stress in
// the fibre frame averaged over the quadrature points (CBElementSolidT10::GetPK2Stress,
// CBElementSolidT10.cpp:817-838, (e1 + e2 + e3 + e4) / 4; T4: its one point) and the
// deformation gradient at the centroid in the global frame (T10: dN/dX at the
// centroid from the reference coordinates, CBElementSolidT10.cpp:307-322 before its
// rotation; T4: the element's constant F).  out[r][E]: r < 9 F, 9 <= r < 18 S (NaN for
// an element with det F <= 0).
template <int NPE, int NQP>
__global__ void __launch_bounds__(128) k_element_export(ElemArgs a, const double* Xref, double* out) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    const int64_t E = a.E;
    const DevMaterial& m = a.mats[a.matidx ? a.matidx[e] : 0];
    double x[3 * NPE];
    for (int k = 0; k < NPE; k++) {
        const int64_t n = a.conn[(int64_t)k * E + e];
        x[3 * k] = a.x[3 * n];
        x[3 * k + 1] = a.x[3 * n + 1];
        x[3 * k + 2] = a.x[3 * n + 2];
    }
    double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool ok = true;
    for (int q = 0; q < NQP; q++) {
        double g[NPE][3];
        load_g<NPE>(a, e, q, g);
        double B[9], Binv[9], Floc[9];
        for (int r = 0; r < 9; r++) B[r] = a.bases[(int64_t)(q * 9 + r) * E + e];
        local_F<NPE>(x, g, B, Binv, Floc);
        Law L;
        if (!eval_law(m, Floc, 0.0, 0.0, false, L)) ok = false;
        for (int r = 0; r < 9; r++) S[r] += L.S[r];
    }
    double F[9];
    if (NPE == 10) {
        double X[30], dN[30];
        for (int k = 0; k < 10; k++) {
            const int64_t n = a.conn[(int64_t)k * E + e];
            for (int c = 0; c < 3; c++) X[3 * k + c] = Xref[3 * n + c];
        }
        t10_shape_point(X, 0.25, 0.25, 0.25, 0.25, dN);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0.0;
                for (int k = 0; k < 10; k++) t += dN[3 * k + j] * x[3 * k + i];
                F[3 * i + j] = t;
            }
    } else {
        double g[NPE][3];
        load_g<NPE>(a, e, 0, g);
        for (int i = 0; i < 3; i++) {
            // CBElementSolidT4::CalcDeformationTensorWithLocalBasis (:130-144), before the rotation
            F[3 * i] = g[0][0] * x[i] + g[1][0] * x[3 + i] + g[2][0] * x[6 + i] + g[3][0] * x[9 + i];
            F[3 * i + 1] = g[0][1] * x[i] + g[1][1] * x[3 + i] + g[2][1] * x[6 + i] + g[3][1] * x[9 + i];
            F[3 * i + 2] = g[0][2] * x[i] + g[1][2] * x[3 + i] + g[2][2] * x[6 + i] + g[3][2] * x[9 + i];
        }
    }
    for (int r = 0; r < 9; r++) {
        out[(int64_t)r * E + e] = F[r];
        out[(int64_t)(9 + r) * E + e] = ok ? S[r] / NQP : NAN;
    }
}

}  // namespace

void launch_element_export(int npe, const ElemArgs& a, const double* Xref, double* out, hipStream_t s) {
    if (a.E == 0) return;
    const int nblk = (int)((a.E + 127) / 128);
    if (npe == 10)
        hipLaunchKernelGGL((k_element_export<10, 4>), dim3(nblk), dim3(128), 0, s, a, Xref, out);
    else
        hipLaunchKernelGGL((k_element_export<4, 1>), dim3(nblk), dim3(128), 0, s, a, Xref, out);
}

void launch_element_residual(int npe, const ElemArgs& a, hipStream_t s) {
    if (a.E == 0) return;
    const int nblk = (int)((a.E + kResidualBlock - 1) / kResidualBlock);
    const bool l17 = a.l17_params != nullptr;
    if (npe == 10 && l17)
        hipLaunchKernelGGL((k_residual<10, 4, true>), dim3(nblk, a.n_sets), dim3(kResidualBlock), 0, s, a);
    else if (npe == 10)
        hipLaunchKernelGGL((k_residual<10, 4>), dim3(nblk, a.n_sets), dim3(kResidualBlock), 0, s, a);
    else if (l17)
        hipLaunchKernelGGL((k_residual<4, 1, true>), dim3(nblk, a.n_sets), dim3(kResidualBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_residual<4, 1>), dim3(nblk, a.n_sets), dim3(kResidualBlock), 0, s, a);
}

void launch_element_residual_tangent(int npe, bool fp32, const ElemArgs& a, hipStream_t s, int variant) {
    if (a.E == 0) return;
    const int nblk = (int)((a.E + kTanElemsPerWave - 1) / kTanElemsPerWave);
    const int nblk2 = (int)((a.E + 2 * kTanElemsPerWave - 1) / (2 * kTanElemsPerWave));
#define CMG_LAUNCH(K) hipLaunchKernelGGL(K, dim3(nblk, a.n_sets), dim3(64), 0, s, a)
#define CMG_LAUNCH2(K) hipLaunchKernelGGL(K, dim3(nblk, a.n_sets), dim3(128), 0, s, a)
#define CMG_LAUNCH_G2(K, T) hipLaunchKernelGGL(K, dim3(nblk2, a.n_sets), dim3(T), 0, s, a)
    if (a.l17_params) {
        // Land17 contexts: one shape per element type (the tuned variants stay free of its code)
        if (npe == 10 && fp32) CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, float, 1, 2, 2, true>), 256);
        else if (npe == 10) CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, double, 1, 2, 2, true>), 256);
        else if (fp32) CMG_LAUNCH((k_residual_tangent_v2<4, 1, float, 1, 1, 1, true>));
        else CMG_LAUNCH((k_residual_tangent_v2<4, 1, double, 1, 1, 1, true>));
    } else if (npe == 10) {
        if (variant == 5) {
            if (fp32) CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, float, 1, 1, 2>), 128);
            else CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, double, 1, 1, 2>), 128);
        } else if (variant == 6) {
            if (fp32) CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, float, 1, 2, 2>), 256);
            else CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, double, 1, 2, 2>), 256);
        } else if (variant == 7) {
            if (fp32) CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, float, 1, 2, 2, false, true>), 256);
            else CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, double, 1, 2, 2, false, true>), 256);
        } else if (variant == 8) {
            if (fp32) CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, float, 1, 1, 2, false, true>), 128);
            else CMG_LAUNCH_G2((k_residual_tangent_v2<10, 4, double, 1, 1, 2, false, true>), 128);
        } else if (variant == 1) {
            if (fp32) CMG_LAUNCH((k_residual_tangent<10, 4, float>));
            else CMG_LAUNCH((k_residual_tangent<10, 4, double>));
        } else if (variant == 3) {
            if (fp32) CMG_LAUNCH((k_residual_tangent_v2<10, 4, float, 2>));
            else CMG_LAUNCH((k_residual_tangent_v2<10, 4, double, 2>));
        } else if (variant == 4) {
            if (fp32) CMG_LAUNCH2((k_residual_tangent_v2<10, 4, float, 1, 2>));
            else CMG_LAUNCH2((k_residual_tangent_v2<10, 4, double, 1, 2>));
        } else {
            if (fp32) CMG_LAUNCH((k_residual_tangent_v2<10, 4, float, 1>));
            else CMG_LAUNCH((k_residual_tangent_v2<10, 4, double, 1>));
        }
    } else {
        if (variant == 1) {
            if (fp32) CMG_LAUNCH((k_residual_tangent<4, 1, float>));
            else CMG_LAUNCH((k_residual_tangent<4, 1, double>));
        } else {
            if (fp32) CMG_LAUNCH((k_residual_tangent_v2<4, 1, float, 1>));
            else CMG_LAUNCH((k_residual_tangent_v2<4, 1, double, 1>));
        }
    }
#undef CMG_LAUNCH
#undef CMG_LAUNCH2
#undef CMG_LAUNCH_G2
}

void launch_surface(const SurfArgs& a, bool tangent, hipStream_t s) {
    if (a.S == 0) return;
    const int nblk = (int)((a.S + 63) / 64);
    if (a.nps == 3)
        hipLaunchKernelGGL((k_surface<3>), dim3(nblk, a.n_sets), dim3(64), 0, s, a, tangent);
    else
        hipLaunchKernelGGL((k_surface<6>), dim3(nblk, a.n_sets), dim3(64), 0, s, a, tangent);
}

void launch_land17_commit(int npe, const ElemArgs& a, double* state, hipStream_t s) {
    if (a.E == 0) return;
    const dim3 grid((unsigned)((a.E + 127) / 128));
    if (npe == 10)
        hipLaunchKernelGGL((k_land17_commit<10, 4>), grid, dim3(128), 0, s, a, state);
    else
        hipLaunchKernelGGL((k_land17_commit<4, 1>), grid, dim3(128), 0, s, a, state);
}

void launch_t10_shape(int64_t E, const int32_t* conn_aos, const double* X, double* dNdX_soa, double* w,
                      hipStream_t s) {
    if (E == 0) return;
    const int nblk = (int)((E + 127) / 128);
    hipLaunchKernelGGL(k_t10_shape, dim3(nblk), dim3(128), 0, s, E, conn_aos, X, dNdX_soa, w);
}

}  // namespace cmg
