// CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates the reference's arithmetic for the mechanics hot path.  Reference
// paths below are relative to the reference checkout's mechanics/src/.
// Templated on the scalar: double for the restatement itself,
// std::complex<double> for complex-step derivatives.

#include "oracle.h"
// data only: the calcium transients of the Land17 model (one copy of the reference's tables)
#include "../cardiomechanics_amd/csrc/land17_calcium.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;

inline double re(double v) { return v; }
inline double re(const cplx& v) { return v.real(); }

// ---------------------------------------------------------------------------
// Row-major 3x3 matrix, restating math_pack::Matrix3 (MathPack/Matrix3.h).
// ---------------------------------------------------------------------------
template <class T> struct M3 {
    T e[9];
};

template <class T> M3<T> zero3() {
    M3<T> r;
    for (int i = 0; i < 9; i++) r.e[i] = T(0.0);
    return r;
}

template <class T> M3<T> from(const double* a) {
    M3<T> r;
    for (int i = 0; i < 9; i++) r.e[i] = T(a[i]);
    return r;
}

// Matrix3::operator* (Matrix3.h:290-303): same term order.
template <class T> M3<T> mul(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
    const T* x = a.e;
    const T* y = b.e;
    r.e[0] = x[0] * y[0] + x[1] * y[3] + x[2] * y[6];
    r.e[1] = x[0] * y[1] + x[1] * y[4] + x[2] * y[7];
    r.e[2] = x[0] * y[2] + x[1] * y[5] + x[2] * y[8];
    r.e[3] = x[3] * y[0] + x[4] * y[3] + x[5] * y[6];
    r.e[4] = x[3] * y[1] + x[4] * y[4] + x[5] * y[7];
    r.e[5] = x[3] * y[2] + x[4] * y[5] + x[5] * y[8];
    r.e[6] = x[6] * y[0] + x[7] * y[3] + x[8] * y[6];
    r.e[7] = x[6] * y[1] + x[7] * y[4] + x[8] * y[7];
    r.e[8] = x[6] * y[2] + x[7] * y[5] + x[8] * y[8];
    return r;
}

template <class T> M3<T> tr(const M3<T>& a) {
    M3<T> r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.e[3 * i + j] = a.e[3 * j + i];
    return r;
}

// Matrix3::Det (Matrix3.h:482-491), Sarrus in the reference's term order.
template <class T> T det(const M3<T>& m) {
    const T* e = m.e;
    return (e[0] * e[4] * e[8]) + (e[3] * e[7] * e[2]) + (e[6] * e[1] * e[5]) - (e[2] * e[4] * e[6]) -
           (e[5] * e[7] * e[0]) - (e[8] * e[3] * e[1]);
}

// Matrix3::GetInverse (Matrix3.h:523-540): adjugate / det, INF on det == 0.
template <class T> M3<T> inv(const M3<T>& m) {
    T d = det(m);
    M3<T> r;
    if (d == T(0.0)) {
        for (int i = 0; i < 9; i++) r.e[i] = T(INFINITY);
        return r;
    }
    const T* e = m.e;
    r.e[0] = (-e[5] * e[7] + e[4] * e[8]) / d;
    r.e[1] = (e[2] * e[7] - e[1] * e[8]) / d;
    r.e[2] = (-e[2] * e[4] + e[1] * e[5]) / d;
    r.e[3] = (e[5] * e[6] - e[3] * e[8]) / d;
    r.e[4] = (-e[2] * e[6] + e[0] * e[8]) / d;
    r.e[5] = (e[2] * e[3] - e[0] * e[5]) / d;
    r.e[6] = (-e[4] * e[6] + e[3] * e[7]) / d;
    r.e[7] = (e[1] * e[6] - e[0] * e[7]) / d;
    r.e[8] = (-e[1] * e[3] + e[0] * e[4]) / d;
    return r;
}

// ---------------------------------------------------------------------------
// Analytic functions, CBData/CBAnalyticFunction.h:53-176.
// ---------------------------------------------------------------------------
double doublehill_raw(const double* fn, double t) {
    double contr = std::pow(t / fn[6], fn[8]);
    contr = contr / (1.0 + contr);
    double relax = std::pow(t / fn[7], fn[9]);
    relax = 1.0 / (1.0 + relax);
    return contr * relax;
}

double doublehill_max(const double* fn) {
    // CBAnalyticFunction.h:155-165: maximum over 1001 samples of one period
    double mx = -INFINITY;
    for (int i = 0; i <= 1000; i++) {
        double t = i * fn[5] / 1000;
        mx = std::max(mx, doublehill_raw(fn, t));
    }
    return mx;
}

double analytic(int type, const double* fn, double time, double prev) {
    const double start = fn[0], stop = fn[1];
    switch (type) {
    case ORC_FN_LINEAR:  // :57-72
        if (time < start) return 0.0;
        if (time < stop + 0.5 * (time - prev)) return (time - start) * 1.0 / (stop - start);
        return fn[2] != 0.0 ? 1.0 : 0.0;
    case ORC_FN_SINUS:  // :86-91
        if ((time >= start) && (time <= stop)) return 0.5 - 0.5 * std::cos(fn[3] * (time - start) + fn[4]);
        return 0.0;
    case ORC_FN_SINUSDRIVER:  // :109-116
        if ((time >= start) && (time <= stop)) return std::sin(3.14159 * (time - start) / (stop - start));
        return 0.0;
    case ORC_FN_DOUBLEHILL: {  // :129-145
        if ((time >= start) && (time <= stop)) {
            double t = time - fn[10];
            if (t > 0.0)
                t = std::fmod(t, fn[5]);
            else
                t = 0.0;
            double dmax = fn[11] > 0.0 ? fn[11] : doublehill_max(fn);
            return doublehill_raw(fn, t) / dmax;
        }
        return 0.0;
    }
    default:
        return 0.0;
    }
}

// Active tension of the FromFunction model (CBTensionFromFunction.cpp:59-61):
// Tmax * g(t - t_act).  Linear is evaluated with the previous time equal to
// the current one (the steady value within a step).
double tension(const orc_material* m, double time, double tact) {
    if (m->tension_fn == ORC_FN_NONE || m->tension_fn == ORC_FN_TANH || m->tension_fn == ORC_FN_LAND17) return 0.0;
    double tt = time - tact;
    return m->tension_max * analytic(m->tension_fn, m->fn, tt, tt);
}

// ---------------------------------------------------------------------------
// Constitutive laws (local fibre frame: e0 = f, e1 = s, e2 = n).
// ---------------------------------------------------------------------------

// CBConstitutiveModelGuccione::CalcPK2Stress, CBConstitutiveModelGuccione.cpp:103-138
template <class T> int pk2_guccione(const orc_material* m, const M3<T>& F, M3<T>& S) {
    const double Cg = m->p[0], bf = m->p[1], bt = m->p[2], bfs = m->p[3], k = m->p[4];
    if (!m->ignore_corrupt) {
        if (re(det(F)) <= 0) return ORC_CORRUPT_ELEMENT;
    }
    T J = det(F);
    M3<T> C = mul(tr(F), F);
    M3<T> E;
    for (int i = 0; i < 9; i++) E.e[i] = 0.5 * (C.e[i] - ((i % 4 == 0) ? 1.0 : 0.0));
    const T* e = E.e;
    M3<T> Cinv = inv(C);
    T Q = bf * e[0] * e[0] + bt * (e[4] * e[4] + e[8] * e[8] + e[5] * e[5] + e[7] * e[7]) +
          bfs * (e[1] * e[1] + e[3] * e[3] + e[2] * e[2] + e[6] * e[6]);
    T expQ = std::exp(Q);
    T* p = S.e;
    p[0] = Cg * bf * e[0] * expQ;
    p[1] = Cg * bfs * e[1] * expQ;
    p[2] = Cg * bfs * e[2] * expQ;
    p[3] = Cg * bfs * e[3] * expQ;
    p[4] = Cg * bt * e[4] * expQ;
    p[5] = Cg * bt * e[5] * expQ;
    p[6] = Cg * bfs * e[6] * expQ;
    p[7] = Cg * bt * e[7] * expQ;
    p[8] = Cg * bt * e[8] * expQ;
    T vol = k * (J - 1.0) * J;
    for (int i = 0; i < 9; i++) p[i] += vol * Cinv.e[i];
    return ORC_SUCCESS;
}

// CBConstitutiveModelGuccione::CalcEnergy, :73-101
template <class T> int energy_guccione(const orc_material* m, const M3<T>& F, T& W) {
    const double Cg = m->p[0], bf = m->p[1], bt = m->p[2], bfs = m->p[3], k = m->p[4];
    if (!m->ignore_corrupt) {
        if (re(det(F)) <= 0) return ORC_CORRUPT_ELEMENT;
    }
    T J = det(F);
    M3<T> C = mul(tr(F), F);
    M3<T> E;
    for (int i = 0; i < 9; i++) E.e[i] = 0.5 * (C.e[i] - ((i % 4 == 0) ? 1.0 : 0.0));
    const T* e = E.e;
    T Q = bf * e[0] * e[0] + bt * (e[4] * e[4] + e[8] * e[8] + e[5] * e[5] + e[7] * e[7]) +
          bfs * (e[1] * e[1] + e[3] * e[3] + e[2] * e[2] + e[6] * e[6]);
    W = Cg / 2. * (std::exp(Q) - 1.0);
    W += k / 2. * (J - 1.0) * (J - 1.0);
    return ORC_SUCCESS;
}

// CBConstitutiveModelHolzapfel::Heavyside, CBConstitutiveModelHolzapfel.cpp:148-154
template <class T> T heaviside(double k, const T& I4) {
    if (k == 0) return (re(I4) >= 1.0) ? T(1.0) : T(0.0);
    return 1.0 / (1.0 + std::exp(-k * (I4 - 1.0)));
}

// CBConstitutiveModelHolzapfel::CalcPK2Stress, :118-146.  Jm = pow(J, -2/3)
// evaluates pow(J, 0) = 1 because -2/3 is integer division (:121).
template <class T> int pk2_holzapfel(const orc_material* m, const M3<T>& F, M3<T>& S) {
    const double a = m->p[0], b = m->p[1], af = m->p[2], bf = m->p[3], as = m->p[4], bs = m->p[5],
                 afs = m->p[6], bfs = m->p[7], kappa = m->p[8], k = m->p[9];
    T J = det(F);
    const double Jm = 1.0;
    if (!m->ignore_corrupt && (re(J) <= 0)) return ORC_CORRUPT_ELEMENT;
    M3<T> C = mul(tr(F), F);
    M3<T> Cinv = inv(C);
    T I1 = C.e[0] + C.e[4] + C.e[8];
    T I4f = C.e[0];
    T I4s = C.e[4];
    T I8fs = 0.5 * (C.e[1] + C.e[3]);
    T cvol = (kappa / 2.0) * (J - 1.0 / J) * J;
    T ciso = Jm * a * std::exp(b * (Jm * I1 - 3.0));
    T cf = 2.0 * af * heaviside(k, I4f) * (I4f - 1.0) * std::exp(bf * (I4f - 1.0) * (I4f - 1.0));
    T cs = 2.0 * as * heaviside(k, I4s) * (I4s - 1.0) * std::exp(bs * (I4s - 1.0) * (I4s - 1.0));
    T cfs = afs * I8fs * std::exp(bfs * I8fs * I8fs);
    for (int i = 0; i < 9; i++) {
        T id = (i % 4 == 0) ? T(1.0) : T(0.0);
        S.e[i] = cvol * Cinv.e[i] + ciso * (id - 1.0 / 3.0 * I1 * Cinv.e[i]);
    }
    S.e[0] += cf;
    S.e[4] += cs;
    S.e[1] += cfs;
    S.e[3] += cfs;
    return ORC_SUCCESS;
}

// CBConstitutiveModelHolzapfel::CalcEnergy, :87-116 (with Jm = 1 as coded)
template <class T> int energy_holzapfel(const orc_material* m, const M3<T>& F, T& W) {
    const double a = m->p[0], b = m->p[1], af = m->p[2], bf = m->p[3], as = m->p[4], bs = m->p[5],
                 afs = m->p[6], bfs = m->p[7], kappa = m->p[8], k = m->p[9];
    T J = det(F);
    const double Jm = 1.0;
    if (!m->ignore_corrupt && (re(J) <= 0)) return ORC_CORRUPT_ELEMENT;
    M3<T> C = mul(tr(F), F);
    T I1 = Jm * (C.e[0] + C.e[4] + C.e[8]);
    T I4f = C.e[0];
    T I4s = C.e[4];
    T I8fs = 0.5 * (C.e[1] + C.e[3]);
    T vol = (kappa / 4) * (J * J - 1.0 - 2.0 * std::log(J));
    T iso = a / (2.0 * b) * (std::exp(b * (I1 - 3.0)) - 1.0);
    T anf = heaviside(k, I4f) * (af / (2.0 * bf) * (std::exp(bf * (I4f - 1.0) * (I4f - 1.0)) - 1.0));
    T ans = heaviside(k, I4s) * (as / (2.0 * bs) * (std::exp(bs * (I4s - 1.0) * (I4s - 1.0)) - 1.0));
    T anfs = afs / (2.0 * bfs) * (std::exp(bfs * I8fs * I8fs) - 1.0);
    W = vol + iso + anf + ans + anfs;
    return ORC_SUCCESS;
}

template <class T> int pk2(const orc_material* m, const M3<T>& F, M3<T>& S) {
    return m->law == ORC_LAW_HOLZAPFEL ? pk2_holzapfel(m, F, S) : pk2_guccione(m, F, S);
}

template <class T> int energy(const orc_material* m, const M3<T>& F, T& W) {
    return m->law == ORC_LAW_HOLZAPFEL ? energy_holzapfel(m, F, W) : energy_guccione(m, F, W);
}

// CBTensionModel::CalcActiveStress, CBTensionModel/CBTensionModel.h:37-50, with
// the FromFunction model's default stress coefficients {1,0,0, 0,0,0, 0,0,0}
// (CBTensionModel.h:100; FromFunction never changes them): only the ff entry
// T / sqrt(I4f) is non-zero, I4f = |F e_f|^2 = |column 0 of F|^2.
template <class T> void active(const M3<T>& F, double Ta, M3<T>& Sa) {
    Sa = zero3<T>();
    T I4f = F.e[0] * F.e[0] + F.e[3] * F.e[3] + F.e[6] * F.e[6];
    Sa.e[0] = Ta * (1. / std::sqrt(I4f));
}

// CBTensionModelTanH::CalcActiveTension (CBTensionModelTanH.cpp:71-97), the
// reference's statement order: lambda = |column 0 of the local F|,
// phi = tanh(ld (lambda - l0)), tau_c = tau_c0 + ld_up (1 - phi), t wrapped into
// one cycle by repeated subtraction, t_s = max(min(t - t_act - t_emd, t_dur), 0),
// Ta = max(S_peak phi tanh(t_s/tau_c)^2 tanh((t_dur - t_s)/tau_r)^2, 0), T = Tmax Ta.
// The activation time includes startTime (AddToActivationTime, :90).
template <class T> T tanh_tension(const orc_material* m, const T& lambda, double time, double tact) {
    const double* f = m->fn;
    const double l0 = f[1], t_emd = f[2], S_peak = f[3], t_dur = f[4], tau_c0 = f[5], ld = f[6], ld_up = f[7];
    const double tau_r = f[8];
    T phi = std::tanh(ld * (lambda - l0));
    T tau_c = tau_c0 + ld_up * (1.0 - phi);
    double t = time;
    while (t > f[0]) t -= f[0];
    const double t_s = std::max(std::min(t - (tact + f[10]) - t_emd, t_dur), 0.0);
    T a = std::tanh(t_s / tau_c);
    const double b = std::tanh((t_dur - t_s) / tau_r);
    T Ta = S_peak * phi * (a * a) * (b * b);
    if (std::real(Ta) < 0.0) Ta = T(0.0);
    return m->tension_max * Ta;
}

// ---------------------------------------------------------------------------
// Land17 (CBTensionModel/CBTensionModelLand17.cpp): one set of cross-bridge states per
// element, advanced by one explicit Euler step from the last committed state at every
// evaluation (CalcActiveTension, :298-372).  Parameters p[40]: TRPNn TRPNk Ca50 ku nTM
// TRPN50 kuw kws rs rw gammas gammaw phi Aeff beta0 beta1 Tref a k eta_l eta_s xi
// (0-21), startTime cycleLength calcium(0 Coppini, 1 Brixius) rateDependancy(0/1) kff
// (22-26), initial XS XW TRPN TmBlocked ZetaS ZetaW Cd (27-33).  State: XS XW TRPN
// TmBlocked ZetaS ZetaW Cd lambda.
// ---------------------------------------------------------------------------
// thread-local inputs of the element being evaluated (set by the batch loops)
struct L17Elem {
    const double* p = nullptr;
    const double* st = nullptr;
    double dt_ms = 0.0;
};
thread_local L17Elem tl_l17;
struct L17Config {
    const double* params = nullptr;   // [n_mats][40]
    const double* state = nullptr;    // [E][8]
    double dt_ms = 0.0;
} g_l17;

// Cai of the element at `time`: CaCoppini / CaBrixius (:55-245) with the activation time
// including startTime (AddToActivationTime, :480)
double land17_cai(const double* p, double time, double tact) {
    double t = time - (tact + p[22]);
    while (t > p[23]) t = t - p[23];
    int index = (int)(t / 1e-3);
    if (index < 0) index = 0;
    index = std::min(1000, index);
    return cmg::kLand17Cai[p[24] != 0.0 ? 1 : 0][index];
}

// one evaluation: tension Tmax * Tension (kPa-scaled by Tmax) at fibre stretch lam; the
// advanced state in next[8] when not null (CalcActiveTension, :298-372; constants of the
// constructor, :262-272)
template <class T>
T land17_tension(const double* p, double tmax, const double* P, const T& lam, double dt, double cai, double* next) {
    const double TRPNn = p[0], TRPNk = p[1], Ca50_ = p[2], ku = p[3], nTM = p[4], TRPN50 = p[5], kuw = p[6];
    const double rs = p[8], rw = p[9], gammas = p[10], gammaw = p[11], phi = p[12], Aeff = p[13], beta0 = p[14];
    const double beta1 = p[15], Tref = p[16], a = p[17], k = p[18], eta_l = p[19], eta_s = p[20], xi = p[21];
    const double kws = xi * p[7];
    const double cw = phi * kuw * ((1.0 - rs) * (1.0 - rw)) / ((1.0 - rs) * rw);
    const double cs = phi * kws * ((1.0 - rs) * rw) / rs;
    const double kwu = kuw * (1.0 / rw - 1.0) - kws;
    const double ksu = (kws * rw * (1.0 / rs - 1.0));
    const double A = Aeff * rs / ((1.0 - rs) * rw + rs);
    const double XSSS = rs * 0.5, XWSS = (1.0 - rs) * rw * 0.5;
    const double ktm_block = ku * std::pow(TRPN50, nTM) * 0.5 / (0.5 - XSSS - XWSS);
    const double XS0 = P[0], XW0 = P[1], TRPN0 = P[2], TB0 = P[3], ZS0 = P[4], ZW0 = P[5], Cd0 = P[6], lam0 = P[7];
    T dldt = T(0.0);
    if (!(dt == 0.0 || p[25] == 0.0)) dldt = (lam - lam0) / dt;
    T lam_m = (std::real(lam) < 1.2) ? lam : T(1.2);
    T lm87 = (std::real(lam_m) < 0.87) ? lam_m : T(0.87);
    T h = 1.0 + beta0 * (lam_m + lm87 - 1.87);
    if (std::real(h) < 0.0) h = T(0.0);
    const double XU = (1.0 - TB0) - XW0 - XS0;
    const double xb_ws = kws * XW0, xb_uw = kuw * XU, xb_wu = kwu * XW0, xb_su = ksu * XS0;
    const double zs_pos = (ZS0 > 0.0) * ZS0;
    const double zs_neg = (ZS0 < -1.0) * (-ZS0 - 1.0);
    const double zs = std::max(zs_pos, zs_neg);
    const double xb_su_gamma = gammas * zs * XS0;
    const double XS = XS0 + dt * (xb_ws - xb_su - xb_su_gamma);
    const double gr_w = (ZW0 < 0.0) ? -ZW0 : ZW0;
    const double xb_wu_gamma = gammaw * gr_w * XW0;
    const double XW = XW0 + dt * (xb_uw - xb_wu - xb_ws - xb_wu_gamma);
    T Ca50 = Ca50_ + beta1 * (lam_m - 1.0);
    T TRPN = TRPN0 + dt * (TRPNk * (std::pow(cai / Ca50, TRPNn) * (1.0 - TRPN0) - TRPN0));
    T ZS = ZS0 + dt * (A * dldt - cs * ZS0);
    T ZW = ZW0 + dt * (A * dldt - cw * ZW0);
    T Ta = h * (Tref / rs) * ((ZS + 1.0) * XS + ZW * XW);
    if (std::real(Ta) < 0.0) Ta = T(0.0);
    T C_s = (lam - 1.0) - Cd0;
    const double eta = (std::real(C_s) > 0.0) ? eta_l : eta_s;
    T F_d = a * k * C_s;
    T Tension = Ta + F_d;
    if (std::real(Tension) < 0.0) Tension = T(0.0);
    if (next) {
        const double TRPNr = std::real(TRPN);
        const double TRPN_NP = std::min(100.0, std::pow(TRPNr, (-nTM / 2.0)));
        next[0] = XS;
        next[1] = XW;
        next[2] = TRPNr;
        next[3] = TB0 + dt * (ktm_block * TRPN_NP * XU - ku * std::pow(TRPNr, nTM / 2.0) * TB0);
        next[4] = std::real(ZS);
        next[5] = std::real(ZW);
        next[6] = Cd0 + dt * (k * std::real(C_s) / eta);
        next[7] = std::real(lam);
    }
    return tmax * Tension;
}

// CalcActiveStress (CBTensionModel.h:37-50) of the element's tension model; TanH and
// Land17 with their default stress coefficients scaled by kff (kss = knn = ksn = 0).
template <class T>
void active_model(const orc_material* m, const M3<T>& F, double time, double tact, M3<T>& Sa) {
    if (m->tension_fn == ORC_FN_LAND17) {
        Sa = zero3<T>();
        const L17Elem& l = tl_l17;
        if (!l.p) return;
        T lam = std::sqrt(F.e[0] * F.e[0] + F.e[3] * F.e[3] + F.e[6] * F.e[6]);
        const double cai = land17_cai(l.p, time, tact);
        Sa.e[0] = l.p[26] * land17_tension<T>(l.p, m->tension_max, l.st, lam, l.dt_ms, cai, nullptr) * (1. / lam);
        return;
    }
    if (m->tension_fn != ORC_FN_TANH) {
        active(F, tension(m, time, tact), Sa);
        return;
    }
    Sa = zero3<T>();
    T I4f = F.e[0] * F.e[0] + F.e[3] * F.e[3] + F.e[6] * F.e[6];
    T lam = std::sqrt(I4f);
    Sa.e[0] = m->fn[9] * tanh_tension(m, lam, time, tact) * (1. / lam);
}

// ---------------------------------------------------------------------------
// T10 element, CBElement/CBElementSolidT10.cpp
// ---------------------------------------------------------------------------

// CalcShapeFunctionDerivatives, :94-264.  Writes 30 values at dNdX, returns detJ.
double t10_shape_point(const double* X, double l1, double l2, double l3, double l4, double* dNdX) {
    const double x1 = X[0], y1 = X[1], z1 = X[2], x2 = X[3], y2 = X[4], z2 = X[5];
    const double x3 = X[6], y3 = X[7], z3 = X[8], x4 = X[9], y4 = X[10], z4 = X[11];
    const double x5 = X[12], y5 = X[13], z5 = X[14], x6 = X[15], y6 = X[16], z6 = X[17];
    const double x7 = X[18], y7 = X[19], z7 = X[20], x8 = X[21], y8 = X[22], z8 = X[23];
    const double x9 = X[24], y9 = X[25], z9 = X[26], x10 = X[27], y10 = X[28], z10 = X[29];
    double Jx1 = 4.0 * (x1 * (l1 - 0.25) + x5 * l2 + x7 * l3 + x8 * l4);
    double Jy1 = 4.0 * (y1 * (l1 - 0.25) + y5 * l2 + y7 * l3 + y8 * l4);
    double Jz1 = 4.0 * (z1 * (l1 - 0.25) + z5 * l2 + z7 * l3 + z8 * l4);
    double Jx2 = 4.0 * (x5 * l1 + x2 * (l2 - 0.25) + x6 * l3 + x9 * l4);
    double Jy2 = 4.0 * (y5 * l1 + y2 * (l2 - 0.25) + y6 * l3 + y9 * l4);
    double Jz2 = 4.0 * (z5 * l1 + z2 * (l2 - 0.25) + z6 * l3 + z9 * l4);
    double Jx3 = 4.0 * (x7 * l1 + x6 * l2 + x3 * (l3 - 0.25) + x10 * l4);
    double Jy3 = 4.0 * (y7 * l1 + y6 * l2 + y3 * (l3 - 0.25) + y10 * l4);
    double Jz3 = 4.0 * (z7 * l1 + z6 * l2 + z3 * (l3 - 0.25) + z10 * l4);
    double Jx4 = 4.0 * (x8 * l1 + x9 * l2 + x10 * l3 + x4 * (l4 - 0.25));
    double Jy4 = 4.0 * (y8 * l1 + y9 * l2 + y10 * l3 + y4 * (l4 - 0.25));
    double Jz4 = 4.0 * (z8 * l1 + z9 * l2 + z10 * l3 + z4 * (l4 - 0.25));
    double Jx12 = Jx1 - Jx2, Jx13 = Jx1 - Jx3, Jx14 = Jx1 - Jx4, Jx23 = Jx2 - Jx3, Jx24 = Jx2 - Jx4, Jx34 = Jx3 - Jx4;
    double Jy12 = Jy1 - Jy2, Jy13 = Jy1 - Jy3, Jy14 = Jy1 - Jy4, Jy23 = Jy2 - Jy3, Jy24 = Jy2 - Jy4, Jy34 = Jy3 - Jy4;
    double Jz12 = Jz1 - Jz2, Jz13 = Jz1 - Jz3, Jz14 = Jz1 - Jz4, Jz23 = Jz2 - Jz3, Jz24 = Jz2 - Jz4, Jz34 = Jz3 - Jz4;
    double Jx21 = -Jx12, Jx31 = -Jx13, Jx32 = -Jx23, Jx42 = -Jx24, Jx43 = -Jx34;
    double Jy21 = -Jy12, Jy31 = -Jy13, Jy32 = -Jy23, Jy42 = -Jy24, Jy43 = -Jy34;
    double Jz21 = -Jz12, Jz31 = -Jz13, Jz32 = -Jz23, Jz42 = -Jz24, Jz43 = -Jz34;
    double detJ = Jx21 * (Jy23 * Jz34 - Jy34 * Jz23) + Jx32 * (Jy34 * Jz12 - Jy12 * Jz34) +
                  Jx43 * (Jy12 * Jz23 - Jy23 * Jz12);
    double a1 = Jy42 * Jz32 - Jy32 * Jz42, a2 = Jy31 * Jz43 - Jy34 * Jz13;
    double a3 = Jy24 * Jz14 - Jy14 * Jz24, a4 = Jy13 * Jz21 - Jy12 * Jz31;
    double b1 = Jx32 * Jz42 - Jx42 * Jz32, b2 = Jx43 * Jz31 - Jx13 * Jz34;
    double b3 = Jx14 * Jz24 - Jx24 * Jz14, b4 = Jx21 * Jz13 - Jx31 * Jz12;
    double c1 = Jx42 * Jy32 - Jx32 * Jy42, c2 = Jx31 * Jy43 - Jx34 * Jy13;
    double c3 = Jx24 * Jy14 - Jx14 * Jy24, c4 = Jx13 * Jy21 - Jx12 * Jy31;
    double factor = 4.0 / detJ;
    const double A[4] = {a1, a2, a3, a4}, B[4] = {b1, b2, b3, b4}, Cc[4] = {c1, c2, c3, c4};
    const double* comp[3] = {A, B, Cc};
    for (int d = 0; d < 3; d++) {
        const double* v = comp[d];
        dNdX[0 + d] = factor * (l1 - 0.25) * v[0];
        dNdX[3 + d] = factor * (l2 - 0.25) * v[1];
        dNdX[6 + d] = factor * (l3 - 0.25) * v[2];
        dNdX[9 + d] = factor * (l4 - 0.25) * v[3];
        dNdX[12 + d] = factor * (l1 * v[1] + l2 * v[0]);
        dNdX[15 + d] = factor * (l2 * v[2] + l3 * v[1]);
        dNdX[18 + d] = factor * (l3 * v[0] + l1 * v[2]);
        dNdX[21 + d] = factor * (l1 * v[3] + l4 * v[0]);
        dNdX[24 + d] = factor * (l2 * v[3] + l4 * v[1]);
        dNdX[27 + d] = factor * (l3 * v[3] + l4 * v[2]);
    }
    return detJ;
}

// CalcDeformationTensorsAtQuadraturePointsWithLocalBasis (:266-288) and the
// residual of CalcNodalForcesHelperFunction (:620-693).
template <class T>
int t10_residual(const T* x, const double* dNdX, const double* bases, double detJ, const uint8_t* bc,
                 const orc_material* m, double time, double tact, T* f) {
    M3<T> Fq[4];
    for (int n = 0; n < 4; n++) {
        const double* g = dNdX + 30 * n;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                T e = T(0.0);
                for (int k = 0; k < 10; k++) e += g[3 * k + j] * x[3 * k + i];
                Fq[n].e[3 * i + j] = e;
            }
    }
    M3<T> Bq[4], BinvT[4];
    for (int n = 0; n < 4; n++) {
        M3<double> Bd = from<double>(bases + 9 * (n + 1));
        M3<double> Bi = tr(inv(Bd));
        Bq[n] = from<T>(Bd.e);
        BinvT[n] = from<T>(Bi.e);
        Fq[n] = mul(mul(tr(Bq[n]), Fq[n]), BinvT[n]);
    }
    M3<T> P[4];
    for (int q = 0; q < 4; q++) {
        M3<T> S;
        int rc = pk2(m, Fq[q], S);
        if (rc != ORC_SUCCESS) return rc;
        M3<T> Sa;
        active_model(m, Fq[q], time, tact, Sa);
        for (int i = 0; i < 9; i++) S.e[i] += Sa.e[i];
        S = tr(mul(Fq[q], S));
        P[q] = mul(mul(BinvT[q], S), tr(Bq[q]));
    }
    double V = detJ / 24.0;
    for (int i = 0; i < 10; i++) {
        T t[3] = {T(0.0), T(0.0), T(0.0)};
        for (int j = 0; j < 4; j++) {
            const double* g = dNdX + 30 * j + 3 * i;
            for (int c = 0; c < 3; c++)
                t[c] += g[0] * P[j].e[0 + c] + g[1] * P[j].e[3 + c] + g[2] * P[j].e[6 + c];
        }
        for (int c = 0; c < 3; c++) f[3 * i + c] = (bc[3 * i + c] == 0) ? V * t[c] : T(0.0);
    }
    return ORC_SUCCESS;
}

template <class T>
int t10_energy(const T* x, const double* dNdX, const double* bases, double detJ, const orc_material* m, T& W) {
    W = T(0.0);
    for (int n = 0; n < 4; n++) {
        const double* g = dNdX + 30 * n;
        M3<T> F;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                T e = T(0.0);
                for (int k = 0; k < 10; k++) e += g[3 * k + j] * x[3 * k + i];
                F.e[3 * i + j] = e;
            }
        M3<double> Bd = from<double>(bases + 9 * (n + 1));
        M3<T> B = from<T>(Bd.e);
        M3<double> Bi = tr(inv(Bd));
        F = mul(mul(tr(B), F), from<T>(Bi.e));
        T w;
        int rc = energy(m, F, w);
        if (rc != ORC_SUCCESS) return rc;
        W += w;
    }
    W *= detJ / 24.0;
    return ORC_SUCCESS;
}

// ---------------------------------------------------------------------------
// T4 element, CBElement/CBElementSolidT4.cpp
// ---------------------------------------------------------------------------
double t4_shape(const double* X, double* dNdX) {  // :73-121
    double z43 = (X[11] - X[8]), z42 = (X[11] - X[5]), z41 = (X[11] - X[2]);
    double z32 = (X[8] - X[5]), z31 = (X[8] - X[2]), z21 = (X[5] - X[2]);
    double y43 = (X[10] - X[7]), y42 = (X[10] - X[4]), y41 = (X[10] - X[1]);
    double y32 = (X[7] - X[4]), y31 = (X[7] - X[1]), y21 = (X[4] - X[1]);
    double x41 = (X[9] - X[0]), x31 = (X[6] - X[0]), x21 = (X[3] - X[0]);
    double d = x21 * (y31 * z41 - y41 * z31) + y21 * (x41 * z31 - x31 * z41) + z21 * (x31 * y41 - x41 * y31);
    dNdX[0] = 1.0 / d * (X[4] * z43 - X[7] * z42 + X[10] * z32);
    dNdX[3] = 1.0 / d * (-X[1] * z43 + X[7] * z41 - X[10] * z31);
    dNdX[6] = 1.0 / d * (X[1] * z42 - X[4] * z41 + X[10] * z21);
    dNdX[9] = 1.0 / d * (-X[1] * z32 + X[4] * z31 - X[7] * z21);
    dNdX[1] = 1.0 / d * (-X[3] * z43 + X[6] * z42 - X[9] * z32);
    dNdX[4] = 1.0 / d * (X[0] * z43 - X[6] * z41 + X[9] * z31);
    dNdX[7] = 1.0 / d * (-X[0] * z42 + X[3] * z41 - X[9] * z21);
    dNdX[10] = 1.0 / d * (X[0] * z32 - X[3] * z31 + X[6] * z21);
    dNdX[2] = 1.0 / d * (X[3] * y43 - X[6] * y42 + X[9] * y32);
    dNdX[5] = 1.0 / d * (-X[0] * y43 + X[6] * y41 - X[9] * y31);
    dNdX[8] = 1.0 / d * (X[0] * y42 - X[3] * y41 + X[9] * y21);
    dNdX[11] = 1.0 / d * (-X[0] * y32 + X[3] * y31 - X[6] * y21);
    return d;
}

// CalcDeformationTensorWithLocalBasis (:130-144) + CalcNodalForcesHelperFunction (:147-210)
template <class T>
int t4_residual(const T* x, const double* dNdX, const double* basis, double vol, const uint8_t* bc,
                const orc_material* m, double time, double tact, T* f) {
    M3<T> F;
    for (int i = 0; i < 3; i++) {
        F.e[i] = dNdX[i] * x[0] + dNdX[i + 3] * x[3] + dNdX[i + 6] * x[6] + dNdX[i + 9] * x[9];
        F.e[3 + i] = dNdX[i] * x[1] + dNdX[i + 3] * x[4] + dNdX[i + 6] * x[7] + dNdX[i + 9] * x[10];
        F.e[6 + i] = dNdX[i] * x[2] + dNdX[i + 3] * x[5] + dNdX[i + 6] * x[8] + dNdX[i + 9] * x[11];
    }
    M3<double> Bd = from<double>(basis);
    M3<T> B = from<T>(Bd.e);
    M3<double> Bi = tr(inv(Bd));
    M3<T> BinvT = from<T>(Bi.e);
    F = mul(mul(tr(B), F), BinvT);
    M3<T> S;
    int rc = pk2(m, F, S);
    if (rc != ORC_SUCCESS) return rc;
    M3<T> Sa;
    active_model(m, F, time, tact, Sa);
    for (int i = 0; i < 9; i++) S.e[i] += Sa.e[i];
    S = mul(S, tr(F));
    S = mul(mul(BinvT, S), tr(B));
    for (int a = 0; a < 4; a++) {
        const double* g = dNdX + 3 * a;
        for (int c = 0; c < 3; c++) {
            int k = 3 * a + c;
            f[k] = (bc[k] == 0) ? vol * (S.e[0 + c] * g[0] + S.e[3 + c] * g[1] + S.e[6 + c] * g[2]) : T(0.0);
        }
    }
    return ORC_SUCCESS;
}

// ---------------------------------------------------------------------------
// Surface loads
// ---------------------------------------------------------------------------
template <class T> void cross(const T* a, const T* b, T* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <class T> T norm3(const T* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// CBElementSurfaceT3::CalcForcesDueToPressure (:257-297) with GetArea (:127-149)
// and GetNormalVector (:99-125) evaluated on the coordinates passed in.  (The
// reference re-reads the unperturbed coordinates from the adapter there, which
// makes its finite-difference pressure tangent identically zero, :268-269.)
template <class T> void t3_forces(const T* x, const uint8_t* bc, double p, T* f) {
    T ab[3], ac[3], n[3];
    for (int i = 0; i < 3; i++) {
        ab[i] = x[i] - x[3 + i];
        ac[i] = x[i] - x[6 + i];
    }
    cross(ab, ac, n);
    T nn = norm3(n);
    T area = 0.5 * nn;
    if (nn != T(0.0))
        for (int i = 0; i < 3; i++) n[i] /= nn;
    T s = -area * p * 1.0 * (1. / 3.);
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) f[3 * a + c] = (bc[3 * a + c] != 0) ? T(0.0) : s * n[c];
}

// Triangle (MathPack/Triangle.h:44-47,135-136): unit normal of (v1-v0)x(v2-v0), area.
template <class T> void tri(const T* v0, const T* v1, const T* v2, T* n, T& area) {
    T a[3], b[3];
    for (int i = 0; i < 3; i++) {
        a[i] = v1[i] - v0[i];
        b[i] = v2[i] - v0[i];
    }
    cross(a, b, n);
    T nn = norm3(n);
    area = 0.5 * nn;
    if (re(area) < 0) area = -area;
    if (nn != T(0.0))
        for (int i = 0; i < 3; i++) n[i] /= nn;
}

template <class T> void normalize(T* v) {
    T nn = norm3(v);
    if (nn != T(0.0))
        for (int i = 0; i < 3; i++) v[i] /= nn;
}

// CBElementSurfaceT6::CalcForcesDueToPressure (:248-285), GetArea (:141-158),
// GetNormalVectorAtQuadraturePoints (:92-122).
template <class T> void t6_forces(const T* x, const uint8_t* bc, double p, T* f) {
    const T *n1 = x, *n2 = x + 3, *n3 = x + 6, *n4 = x + 9, *n5 = x + 12, *n6 = x + 15;
    T t1n[3], t2n[3], t3n[3], t4n[3], A1, A2, A3, A4;
    tri(n1, n4, n6, t1n, A1);
    tri(n4, n5, n6, t2n, A2);
    tri(n4, n2, n5, t3n, A3);
    tri(n6, n5, n3, t4n, A4);
    T area = A1 + A2 + A3 + A4;
    T nv1[3], nv2[3], nv3[3];
    for (int i = 0; i < 3; i++) {
        nv1[i] = t1n[i] + t2n[i];
        nv2[i] = t3n[i] + t2n[i];
        nv3[i] = t4n[i] + t2n[i];
    }
    normalize(nv1);
    normalize(nv2);
    normalize(nv3);
    const double q[3][3] = {{2. / 3., 1. / 6., 1. / 6.}, {1. / 6., 2. / 3., 1. / 6.}, {1. / 6., 1. / 6., 2. / 3.}};
    double N[6][3];
    for (int k = 0; k < 3; k++) {
        double l1 = q[k][0], l2 = q[k][1], l3 = q[k][2];
        N[0][k] = l1 * (2 * l1 - 1);
        N[1][k] = l2 * (2 * l2 - 1);
        N[2][k] = l3 * (2 * l3 - 1);
        N[3][k] = 4 * l1 * l2;
        N[4][k] = 4 * l1 * l3;
        N[5][k] = 4 * l2 * l3;
    }
    T s = -area * p * 1.0 / 3.0;
    for (int i = 0; i < 6; i++)
        for (int c = 0; c < 3; c++) {
            T v = N[i][0] * nv1[c] + N[i][1] * nv2[c] + N[i][2] * nv3[c];
            f[3 * i + c] = (bc[3 * i + c] != 0) ? T(0.0) : s * v;
        }
}

}  // namespace

// ===========================================================================
// extern "C" API
// ===========================================================================
extern "C" {

double orc_analytic(int32_t type, const double* fn, double time, double prev_time) {
    return analytic(type, fn, time, prev_time);
}

int32_t orc_pk2(const orc_material* m, const double* F, double* S) {
    M3<double> Fm = from<double>(F), Sm;
    int rc = pk2(m, Fm, Sm);
    if (rc == ORC_SUCCESS) std::memcpy(S, Sm.e, sizeof(Sm.e));
    return rc;
}

int32_t orc_energy(const orc_material* m, const double* F, double* W) {
    M3<double> Fm = from<double>(F);
    double w = 0;
    int rc = energy(m, Fm, w);
    *W = w;
    return rc;
}

static const double kCsStep = 1e-30;

int32_t orc_energy_dF_cs(const orc_material* m, const double* F, double* dWdF) {
    for (int b = 0; b < 9; b++) {
        M3<cplx> Fc = from<cplx>(F);
        Fc.e[b] += cplx(0.0, kCsStep);
        cplx w;
        int rc = energy(m, Fc, w);
        if (rc != ORC_SUCCESS) return rc;
        dWdF[b] = w.imag() / kCsStep;
    }
    return ORC_SUCCESS;
}

int32_t orc_pk2_dF_cs(const orc_material* m, const double* F, double* dSdF) {
    for (int b = 0; b < 9; b++) {
        M3<cplx> Fc = from<cplx>(F), S;
        Fc.e[b] += cplx(0.0, kCsStep);
        int rc = pk2(m, Fc, S);
        if (rc != ORC_SUCCESS) return rc;
        for (int a = 0; a < 9; a++) dSdF[9 * a + b] = S.e[a].imag() / kCsStep;
    }
    return ORC_SUCCESS;
}

double orc_tanh_tension(const orc_material* m, double lambda, double time, double tact) {
    return tanh_tension<double>(m, lambda, time, tact);
}

void orc_active_stress_model(const orc_material* m, const double* F, double time, double tact, double* Sa) {
    M3<double> S;
    active_model<double>(m, from<double>(F), time, tact, S);
    std::memcpy(Sa, S.e, sizeof(S.e));
}

void orc_active_stress(const double* F, double tension_value, double* Sa) {
    M3<double> Fm = from<double>(F), S;
    active(Fm, tension_value, S);
    std::memcpy(Sa, S.e, sizeof(S.e));
}

// UpdateShapeFunctions -> CalcShapeFunctionDerivativesAtQuadraturePoints (:79-89):
// 4 QPs a=(5+3 sqrt5)/20, b=(5-sqrt5)/20, then the centroid; detJ_ is the last (centroid) value.
void orc_t10_shape(const double* X30, double* dNdX150, double* detJ) {
    const double a = (5 + 3 * std::sqrt(5)) / 20, b = (5 - std::sqrt(5)) / 20;
    t10_shape_point(X30, a, b, b, b, dNdX150 + 0);
    t10_shape_point(X30, b, a, b, b, dNdX150 + 30);
    t10_shape_point(X30, b, b, a, b, dNdX150 + 60);
    t10_shape_point(X30, b, b, b, a, dNdX150 + 90);
    *detJ = t10_shape_point(X30, 0.25, 0.25, 0.25, 0.25, dNdX150 + 120);
}

int32_t orc_t10_residual(const double* x30, const double* dNdX150, const double* bases45, double detJ,
                         const uint8_t* bc30, const orc_material* m, double time, double tact, double* f30) {
    return t10_residual<double>(x30, dNdX150, bases45, detJ, bc30, m, time, tact, f30);
}

// CalcNodalForcesJacobian (:511-567): central differences, 30*i + 3k + c holds
// d f_{3k+c} / d x_i (row of the assembled matrix = perturbed DOF i).
int32_t orc_t10_fd_tangent(const double* x30, const double* dNdX150, const double* bases45, double detJ,
                           const uint8_t* bc30, const orc_material* m, double time, double tact, double eps,
                           double* K900) {
    double x[30], f1[30], f2[30];
    std::memcpy(x, x30, sizeof(x));
    for (int i = 0; i < 900; i++) K900[i] = 0.0;
    double eps2 = 2.0 * eps;
    for (int i = 0; i < 30; i++) {
        if (bc30[i]) continue;
        double xi = x[i];
        x[i] = xi + eps;
        int rc = t10_residual<double>(x, dNdX150, bases45, detJ, bc30, m, time, tact, f1);
        if (rc != ORC_SUCCESS) return rc;
        x[i] = xi - eps;
        rc = t10_residual<double>(x, dNdX150, bases45, detJ, bc30, m, time, tact, f2);
        x[i] = xi;
        if (rc != ORC_SUCCESS) return rc;
        for (int k = 0; k < 30; k++) K900[30 * i + k] = (f1[k] - f2[k]) / eps2;
    }
    return ORC_SUCCESS;
}

// Exact Jacobian by complex step, standard layout K[30*r + c] = d f_r / d x_c,
// with the columns of constrained DOFs left zero like the reference's tangent.
int32_t orc_t10_cs_tangent(const double* x30, const double* dNdX150, const double* bases45, double detJ,
                           const uint8_t* bc30, const orc_material* m, double time, double tact, double* K900) {
    cplx x[30], f[30];
    for (int i = 0; i < 900; i++) K900[i] = 0.0;
    for (int c = 0; c < 30; c++) {
        if (bc30[c]) continue;
        for (int i = 0; i < 30; i++) x[i] = cplx(x30[i], 0.0);
        x[c] += cplx(0.0, kCsStep);
        int rc = t10_residual<cplx>(x, dNdX150, bases45, detJ, bc30, m, time, tact, f);
        if (rc != ORC_SUCCESS) return rc;
        for (int r = 0; r < 30; r++) K900[30 * r + c] = f[r].imag() / kCsStep;
    }
    return ORC_SUCCESS;
}

// CalcConsistentMassMatrix (:352-394)
void orc_t10_mass_consistent(double detJ, double density, const uint8_t* bc30, double* M900) {
    static const double r1[100] = {6,  1,  1,  1,  -4, -6, -4, -4, -6, -6, 1,  6,  1,  1,  -4, -4, -6, -6, -4, -6,
                                   1,  1,  6,  1,  -6, -4, -4, -6, -6, -4, 1,  1,  1,  6,  -6, -6, -6, -4, -4, -4,
                                   -4, -4, -6, -6, 32, 16, 16, 16, 16, 8,  -6, -4, -4, -6, 16, 32, 16, 8,  16, 16,
                                   -4, -6, -4, -6, 16, 16, 32, 16, 8,  16, -4, -6, -6, -4, 16, 8,  16, 32, 16, 16,
                                   -6, -4, -6, -4, 16, 16, 8,  16, 32, 16, -6, -6, -4, -4, 8,  16, 16, 16, 16, 32};
    double c = (density * detJ) / 2520;
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 10; j++)
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++)
                    M900[30 * (3 * i + k) + (3 * j + l)] =
                        ((k == l) && (bc30[3 * i + k] == 0)) ? r1[10 * i + j] * c : 0.0;
}

// CalcLumpedMassMatrix (:396-422)
void orc_t10_mass_lumped(double volume, double density, const uint8_t* bc30, double* M900) {
    double c = (density * volume) / 10;
    for (int i = 0; i < 30; i++)
        for (int j = 0; j < 30; j++) M900[30 * i + j] = ((i == j) && (bc30[i] == 0)) ? c : 0.0;
}

int32_t orc_t10_energy(const double* x30, const double* dNdX150, const double* bases45, double detJ,
                       const orc_material* m, double* W) {
    double w = 0;
    int rc = t10_energy<double>(x30, dNdX150, bases45, detJ, m, w);
    *W = w;
    return rc;
}

int32_t orc_t10_energy_grad_cs(const double* x30, const double* dNdX150, const double* bases45, double detJ,
                               const orc_material* m, double* g30) {
    cplx x[30];
    for (int c = 0; c < 30; c++) {
        for (int i = 0; i < 30; i++) x[i] = cplx(x30[i], 0.0);
        x[c] += cplx(0.0, kCsStep);
        cplx w;
        int rc = t10_energy<cplx>(x, dNdX150, bases45, detJ, m, w);
        if (rc != ORC_SUCCESS) return rc;
        g30[c] = w.imag() / kCsStep;
    }
    return ORC_SUCCESS;
}

void orc_t4_shape(const double* X12, double* dNdX12, double* detJ) { *detJ = t4_shape(X12, dNdX12); }

int32_t orc_t4_residual(const double* x12, const double* dNdX12, const double* basis9, double volume,
                        const uint8_t* bc12, const orc_material* m, double time, double tact, double* f12) {
    return t4_residual<double>(x12, dNdX12, basis9, volume, bc12, m, time, tact, f12);
}

// CBElementSolidT4::CalcNodalForcesJacobian (:294-351): K[12*k + i] = d f_k / d x_i (not transposed)
int32_t orc_t4_fd_tangent(const double* x12, const double* dNdX12, const double* basis9, double volume,
                          const uint8_t* bc12, const orc_material* m, double time, double tact, double eps,
                          double* K144) {
    double x[12], f1[12], f2[12];
    std::memcpy(x, x12, sizeof(x));
    for (int i = 0; i < 144; i++) K144[i] = 0.0;
    double eps2 = 2.0 * eps;
    for (int i = 0; i < 12; i++) {
        if (bc12[i]) continue;
        double xi = x[i];
        x[i] = xi + eps;
        int rc = t4_residual<double>(x, dNdX12, basis9, volume, bc12, m, time, tact, f1);
        if (rc != ORC_SUCCESS) return rc;
        x[i] = xi - eps;
        rc = t4_residual<double>(x, dNdX12, basis9, volume, bc12, m, time, tact, f2);
        x[i] = xi;
        if (rc != ORC_SUCCESS) return rc;
        for (int k = 0; k < 12; k++) K144[12 * k + i] = bc12[k] ? 0.0 : (f1[k] - f2[k]) / eps2;
    }
    return ORC_SUCCESS;
}

int32_t orc_t4_cs_tangent(const double* x12, const double* dNdX12, const double* basis9, double volume,
                          const uint8_t* bc12, const orc_material* m, double time, double tact, double* K144) {
    cplx x[12], f[12];
    for (int i = 0; i < 144; i++) K144[i] = 0.0;
    for (int c = 0; c < 12; c++) {
        if (bc12[c]) continue;
        for (int i = 0; i < 12; i++) x[i] = cplx(x12[i], 0.0);
        x[c] += cplx(0.0, kCsStep);
        int rc = t4_residual<cplx>(x, dNdX12, basis9, volume, bc12, m, time, tact, f);
        if (rc != ORC_SUCCESS) return rc;
        for (int r = 0; r < 12; r++) K144[12 * r + c] = f[r].imag() / kCsStep;
    }
    return ORC_SUCCESS;
}

// CBElementSolidT4::CalcConsistentMassMatrix (:365-410)
void orc_t4_mass_consistent(double volume, double density, const uint8_t* bc12, double* M144) {
    double c = (density * volume) / 20;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++)
                    M144[12 * (3 * i + k) + (3 * j + l)] =
                        ((k == l) && (bc12[3 * i + k] == 0)) ? ((i == j) ? 2.0 : 1.0) * c : 0.0;
}

// CBElementSolidT4 lumped mass (CBElementSolidT4.cpp:432-456): rho V / 4 on every
// diagonal entry (the row sums of the consistent T4 mass), rows of constrained DOFs zero
void orc_t4_mass_lumped(double volume, double density, const uint8_t* bc12, double* M144) {
    double c = (density * volume) / 4;
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) M144[12 * i + j] = ((i == j) && (bc12[i] == 0)) ? c : 0.0;
}

void orc_t3_pressure_forces(const double* x9, const uint8_t* bc9, double p, double* f9) {
    t3_forces<double>(x9, bc9, p, f9);
}

void orc_t3_pressure_cs_tangent(const double* x9, const uint8_t* bc9, double p, double* K81) {
    cplx x[9], f[9];
    for (int c = 0; c < 9; c++) {
        for (int i = 0; i < 9; i++) x[i] = cplx(x9[i], 0.0);
        x[c] += cplx(0.0, kCsStep);
        t3_forces<cplx>(x, bc9, p, f);
        for (int r = 0; r < 9; r++) K81[9 * r + c] = bc9[c] ? 0.0 : f[r].imag() / kCsStep;
    }
}

// CBElementSurfaceT3::CalcContributionToVolume (:151-191)
double orc_t3_volume(const double* x, const double* r) {
    double a[3] = {x[3] - x[0], x[4] - x[1], x[5] - x[2]};
    double b[3] = {x[6] - x[0], x[7] - x[1], x[8] - x[2]};
    double c[3];
    cross(a, b, c);
    double d[3] = {x[0] - r[0], x[1] - r[1], x[2] - r[2]};
    return (d[0] * c[0] + d[1] * c[1] + d[2] * c[2]) / 6.0;
}

void orc_t6_pressure_forces(const double* x18, const uint8_t* bc18, double p, double* f18) {
    t6_forces<double>(x18, bc18, p, f18);
}

// CBElementSurfaceT6::CalcPressureJacobian (:321-366): K[18*(3k+l) + (3i+j)]
void orc_t6_pressure_fd_tangent(const double* x18, const uint8_t* bc18, double p, double eps, double* K324) {
    double x[18], f1[18], f2[18];
    std::memcpy(x, x18, sizeof(x));
    for (int c = 0; c < 18; c++) {
        double tmp = x[c];
        x[c] += eps;
        t6_forces<double>(x, bc18, p, f1);
        x[c] -= 2.0 * eps;
        t6_forces<double>(x, bc18, p, f2);
        for (int r = 0; r < 18; r++)
            K324[18 * r + c] = (bc18[r] == 0 && bc18[c] == 0) ? (f1[r] - f2[r]) / (2 * eps) : 0.0;
        x[c] = tmp;
    }
}

void orc_t6_pressure_cs_tangent(const double* x18, const uint8_t* bc18, double p, double* K324) {
    cplx x[18], f[18];
    for (int c = 0; c < 18; c++) {
        for (int i = 0; i < 18; i++) x[i] = cplx(x18[i], 0.0);
        x[c] += cplx(0.0, kCsStep);
        t6_forces<cplx>(x, bc18, p, f);
        for (int r = 0; r < 18; r++) K324[18 * r + c] = bc18[c] ? 0.0 : f[r].imag() / kCsStep;
    }
}

// CBElementSurfaceT6::CalcContributionToVolume (:191-241)
double orc_t6_volume(const double* x, const double* r) {
    double a1[3], a2[3], a3[3], a4[3], a5[3];
    for (int i = 0; i < 3; i++) {
        a1[i] = x[9 + i] - x[i];
        a2[i] = x[3 + i] - x[i];
        a3[i] = x[12 + i] - x[i];
        a4[i] = x[6 + i] - x[i];
        a5[i] = x[15 + i] - x[i];
    }
    double c1[3], c2[3], c3[3], c4[3];
    cross(a1, a2, c1);
    cross(a2, a3, c2);
    cross(a3, a4, c3);
    cross(a4, a5, c4);
    double d[3] = {x[0] - r[0], x[1] - r[1], x[2] - r[2]};
    double v1 = d[0] * c1[0] + d[1] * c1[1] + d[2] * c1[2];
    double v2 = d[0] * c2[0] + d[1] * c2[1] + d[2] * c2[2];
    double v3 = d[0] * c3[0] + d[1] * c3[1] + d[2] * c3[2];
    double v4 = d[0] * c4[0] + d[1] * c4[1] + d[2] * c4[2];
    return (v1 + v2 + v3 + v4) / 6.0;
}

// ---------------------------------------------------------------------------
// Mesh-wide batches: gather coordinates and BC flags, one element per task.
// ---------------------------------------------------------------------------
// the Land17 inputs of element e (material mi) for this thread, when configured
static void set_l17(int32_t mi, int64_t e) {
    tl_l17 = L17Elem();
    if (g_l17.params) {
        tl_l17.p = g_l17.params + 40 * (int64_t)mi;
        tl_l17.st = g_l17.state + 8 * e;
        tl_l17.dt_ms = g_l17.dt_ms;
    }
}

// fibre-frame deformation gradient of element e at quadrature point q (T10: q 0..3 of
// :266-288; T4: :130-144), for the Land17 state commit
static void local_F_at(int npe, const double* x, const double* dNdX, const double* bases, int q, double* F9) {
    M3<double> F;
    if (npe == 10) {
        const double* g = dNdX + 30 * q;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double v = 0.0;
                for (int k = 0; k < 10; k++) v += g[3 * k + j] * x[3 * k + i];
                F.e[3 * i + j] = v;
            }
    } else {
        for (int i = 0; i < 3; i++) {
            F.e[i] = dNdX[i] * x[0] + dNdX[i + 3] * x[3] + dNdX[i + 6] * x[6] + dNdX[i + 9] * x[9];
            F.e[3 + i] = dNdX[i] * x[1] + dNdX[i + 3] * x[4] + dNdX[i + 6] * x[7] + dNdX[i + 9] * x[10];
            F.e[6 + i] = dNdX[i] * x[2] + dNdX[i + 3] * x[5] + dNdX[i + 6] * x[8] + dNdX[i + 9] * x[11];
        }
    }
    M3<double> Bd = from<double>(bases + 9 * (npe == 10 ? q + 1 : 0));
    M3<double> Bi = tr(inv(Bd));
    F = mul(mul(tr(Bd), F), Bi);
    std::memcpy(F9, F.e, sizeof(F.e));
}

static void gather(int npe, const int32_t* conn, const double* coords, const uint8_t* bcmask, double* x,
                   uint8_t* bc) {
    for (int a = 0; a < npe; a++) {
        int n = conn[a];
        for (int c = 0; c < 3; c++) {
            x[3 * a + c] = coords[3 * (int64_t)n + c];
            bc[3 * a + c] = (bcmask[n] >> c) & 1;
        }
    }
}

int32_t orc_residual_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                           const double* dNdX, const double* bases, const double* w, const uint8_t* bcmask,
                           const orc_material* mats, const int32_t* matidx, const double* tact, double time,
                           double* fe, int32_t* status) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < n_elem; e++) {
        double x[30];
        uint8_t bc[30];
        gather(npe, conn + e * npe, coords, bcmask, x, bc);
        const orc_material* m = mats + (matidx ? matidx[e] : 0);
        double ta = tact ? tact[e] : 0.0;
        set_l17(matidx ? matidx[e] : 0, e);
        int rc;
        if (npe == 10)
            rc = t10_residual<double>(x, dNdX + 150 * e, bases + 45 * e, w[e], bc, m, time, ta, fe + 30 * e);
        else
            rc = t4_residual<double>(x, dNdX + 12 * e, bases + 9 * e, w[e], bc, m, time, ta, fe + 12 * e);
        status[e] = rc;
        bad += (rc != ORC_SUCCESS);
    }
    return bad ? ORC_FAILED : ORC_SUCCESS;
}

int32_t orc_fd_tangent_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                             const double* dNdX, const double* bases, const double* w, const uint8_t* bcmask,
                             const orc_material* mats, const int32_t* matidx, const double* tact, double time,
                             double eps, double* Ke, int32_t* status) {
    int bad = 0;
    const int nd = 3 * npe;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < n_elem; e++) {
        double x[30];
        uint8_t bc[30];
        gather(npe, conn + e * npe, coords, bcmask, x, bc);
        const orc_material* m = mats + (matidx ? matidx[e] : 0);
        double ta = tact ? tact[e] : 0.0;
        set_l17(matidx ? matidx[e] : 0, e);
        int rc;
        if (npe == 10)
            rc = orc_t10_fd_tangent(x, dNdX + 150 * e, bases + 45 * e, w[e], bc, m, time, ta, eps,
                                    Ke + (int64_t)nd * nd * e);
        else
            rc = orc_t4_fd_tangent(x, dNdX + 12 * e, bases + 9 * e, w[e], bc, m, time, ta, eps,
                                   Ke + (int64_t)nd * nd * e);
        status[e] = rc;
        bad += (rc != ORC_SUCCESS);
    }
    return bad ? ORC_FAILED : ORC_SUCCESS;
}

int32_t orc_cs_tangent_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                             const double* dNdX, const double* bases, const double* w, const uint8_t* bcmask,
                             const orc_material* mats, const int32_t* matidx, const double* tact, double time,
                             double* Ke, int32_t* status) {
    int bad = 0;
    const int nd = 3 * npe;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < n_elem; e++) {
        double x[30];
        uint8_t bc[30];
        gather(npe, conn + e * npe, coords, bcmask, x, bc);
        const orc_material* m = mats + (matidx ? matidx[e] : 0);
        double ta = tact ? tact[e] : 0.0;
        set_l17(matidx ? matidx[e] : 0, e);
        int rc;
        if (npe == 10)
            rc = orc_t10_cs_tangent(x, dNdX + 150 * e, bases + 45 * e, w[e], bc, m, time, ta,
                                    Ke + (int64_t)nd * nd * e);
        else
            rc = orc_t4_cs_tangent(x, dNdX + 12 * e, bases + 9 * e, w[e], bc, m, time, ta,
                                   Ke + (int64_t)nd * nd * e);
        status[e] = rc;
        bad += (rc != ORC_SUCCESS);
    }
    return bad ? ORC_FAILED : ORC_SUCCESS;
}

// shape-function derivatives of every element (T10: orc_t10_shape; T4: orc_t4_shape),
// dNdX [E][150] / [E][12] and detJ [E] (CBElementSolid*::CalcShapeFunctionDerivatives)
void orc_shape_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords, double* dNdX,
                     double* detJ) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_elem; e++) {
        double X[30];
        for (int a = 0; a < npe; a++)
            for (int c = 0; c < 3; c++) X[3 * a + c] = coords[3 * (int64_t)conn[e * npe + a] + c];
        if (npe == 10)
            orc_t10_shape(X, dNdX + 150 * e, detJ + e);
        else
            orc_t4_shape(X, dNdX + 12 * e, detJ + e);
    }
}

// element mass matrices of every element: T10 consistent (w = detJ) or lumped (w = detJ,
// volume detJ/6), T4 consistent (w = volume); rows of constrained DOFs zero
void orc_mass_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* w, const double* density,
                    const uint8_t* bcmask, int32_t consistent, double* Me) {
    const int nd = 3 * npe;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_elem; e++) {
        uint8_t bc[30];
        for (int a = 0; a < npe; a++)
            for (int c = 0; c < 3; c++) bc[3 * a + c] = (bcmask[conn[e * npe + a]] >> c) & 1;
        double* out = Me + (int64_t)nd * nd * e;
        if (npe == 10) {
            if (consistent)
                orc_t10_mass_consistent(w[e], density[e], bc, out);
            else
                orc_t10_mass_lumped(w[e] / 6, density[e], bc, out);
        } else {
            if (consistent)
                orc_t4_mass_consistent(w[e], density[e], bc, out);
            else
                orc_t4_mass_lumped(w[e], density[e], bc, out);
        }
    }
}

// Land17: the per-element inputs used by the element batches from now on (params
// [n_mats][40], state [E][8], dt_ms = (time - time of the committed state) * 1000);
// params NULL switches it off
void orc_land17_configure(const double* params, const double* state, double dt_ms) {
    g_l17.params = params;
    g_l17.state = state;
    g_l17.dt_ms = dt_ms;
}

double orc_land17_cai(const double* p, double time, double tact) { return land17_cai(p, time, tact); }

// one Land17 evaluation with an explicit Cai: tension (Tmax-scaled), next state [8] (may be
// NULL), dT/dlambda by complex step
double orc_land17_step(const double* p, double tmax, const double* state, double lambda, double dt_ms, double cai,
                       double* next, double* dTdl) {
    if (dTdl) {
        const cplx t = land17_tension<cplx>(p, tmax, state, cplx(lambda, kCsStep), dt_ms, cai, nullptr);
        *dTdl = t.imag() / kCsStep;
    }
    return land17_tension<double>(p, tmax, state, lambda, dt_ms, cai, next);
}

// the state commit of an accepted step: every Land17 element advances its state from the
// evaluation at its last quadrature point (the last CalcActiveTension call of the step,
// SaveStateVariablesAsNeeded :274-293)
void orc_land17_commit_batch(int32_t npe, int64_t n_elem, const int32_t* conn, const double* coords,
                             const double* dNdX, const double* bases, const orc_material* mats,
                             const int32_t* matidx, const double* tact, const double* params, double time,
                             double dt_ms, double* state) {
    const int nq = npe == 10 ? 4 : 1;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_elem; e++) {
        const int32_t mi = matidx ? matidx[e] : 0;
        if (mats[mi].tension_fn != ORC_FN_LAND17) continue;
        double x[30] = {}, F[9];
        for (int a = 0; a < npe; a++)
            for (int c = 0; c < 3; c++) x[3 * a + c] = coords[3 * (int64_t)conn[e * npe + a] + c];
        local_F_at(npe, x, dNdX + (npe == 10 ? 150 : 12) * e, bases + (npe == 10 ? 45 : 9) * e, nq - 1, F);
        const double lam = std::sqrt(F[0] * F[0] + F[3] * F[3] + F[6] * F[6]);
        const double* p = params + 40 * (int64_t)mi;
        double next[8];
        land17_tension<double>(p, mats[mi].tension_max, state + 8 * e, lam, dt_ms,
                               land17_cai(p, time, tact ? tact[e] : 0.0), next);
        std::memcpy(state + 8 * e, next, sizeof(next));
    }
}

int32_t orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
