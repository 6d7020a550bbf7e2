// physics.h -- Euler / Navier-Stokes point physics, written once for host and device.
//
// Every function is a template on the arithmetic type T (double for the FP64 path,
// float for the mixed path) and is instantiated both in the device kernels and in the
// host plugin functions that euler_convective / euler_boundary / navierstokes_viscous
// return.  Operation order and association follow the reference exactly so that the
// FP64 instantiation is bit-identical to it when compiled with -ffp-contract=off:
//   src/euler/convective.c:15-591, src/euler/boundary.c:7-271, src/euler/create.c:6-44,
//   src/euler/riemann.c:1-197, src/navierstokes/viscous.c:11-39,
//   src/navierstokes/boundary.c:6-18, src/navierstokes/create.c:7-26.
// State layout: primitive {rho, u, v, w, p}; conserved {rho, rho u, rho v, rho w, E};
// basis is row-major {n, t1, t2} (9 values), n = face normal (left -> right).
#pragma once

#include <hip/hip_runtime.h>

#define TEAL_HD __host__ __device__ inline

namespace fv {

enum FluxKind {
    FLUX_NONE = 0,
    FLUX_GODUNOV,
    FLUX_ROE,
    FLUX_HLL,
    FLUX_HLLC,
    FLUX_HLLE,
    FLUX_AUSMD,
    FLUX_AUSMDV,
    FLUX_LXF,
    FLUX_CENTRAL,
    NUM_FLUX
};

enum BcKind {
    BC_NONE = 0,
    BC_SYMMETRY,
    BC_SUPERSONIC_INFLOW,
    BC_SUPERSONIC_OUTFLOW,
    BC_SUBSONIC_INFLOW,
    BC_SUBSONIC_OUTFLOW,
    BC_PRESSURE_OUTFLOW,
    BC_FARFIELD,
    BC_ADIABATIC_WALL,
    NUM_BC
};

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T> TEAL_HD V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> TEAL_HD V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> TEAL_HD V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> TEAL_HD V3<T> mul(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> TEAL_HD V3<T> divv(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> TEAL_HD T dot(V3<T> a, V3<T> b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
template <typename T> TEAL_HD T sqr(T a) { return a * a; }

// row-major 3x3 (src/core/matrix.c:7-33)
template <typename T>
struct M3 {
    V3<T> x, y, z;
};
template <typename T> TEAL_HD M3<T> m3(const T *b)
{
    return M3<T>{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}, {b[6], b[7], b[8]}};
}
template <typename T> TEAL_HD V3<T> mv(M3<T> m, V3<T> v) { return {dot(m.x, v), dot(m.y, v), dot(m.z, v)}; }
template <typename T> TEAL_HD M3<T> tr(M3<T> m)
{
    return M3<T>{{m.x.x, m.y.x, m.z.x}, {m.x.y, m.y.y, m.z.y}, {m.x.z, m.y.z, m.z.z}};
}

template <typename T> TEAL_HD T tmin(T a, T b) { return fmin(a, b); }
template <typename T> TEAL_HD T tmax(T a, T b) { return fmax(a, b); }
template <typename T> TEAL_HD T tabs(T a) { return fabs(a); }
template <typename T> TEAL_HD T tsqrt(T a) { return sqrt(a); }
template <> TEAL_HD float tmin<float>(float a, float b) { return fminf(a, b); }
template <> TEAL_HD float tmax<float>(float a, float b) { return fmaxf(a, b); }
template <> TEAL_HD float tabs<float>(float a) { return fabsf(a); }
template <> TEAL_HD float tsqrt<float>(float a) { return sqrtf(a); }

// face-frame state (src/euler/convective.c:7-13)
template <typename T>
struct Local {
    T rho;
    V3<T> u;
    T p;
    V3<T> m;
    T e;
};

template <typename T>
struct Cons {
    T rho;
    V3<T> m;
    T e;
};

// src/euler/convective.c:15-25
template <typename T> TEAL_HD Local<T> to_local(const T *g, T gamma, M3<T> b)
{
    Local<T> l;
    l.rho = g[0];
    l.u = mv(b, v3(g[1], g[2], g[3]));
    l.p = g[4];
    l.m = mul(l.rho, l.u);
    l.e = (l.p / (gamma - 1)) + (dot(l.m, l.u) / 2);
    return l;
}

// src/euler/convective.c:27-30
template <typename T> TEAL_HD void to_global(T *flux, Cons<T> f, M3<T> b)
{
    V3<T> m = mv(tr(b), f.m);
    flux[0] = f.rho;
    flux[1] = m.x;
    flux[2] = m.y;
    flux[3] = m.z;
    flux[4] = f.e;
}

// src/euler/convective.c:32-41
template <typename T> TEAL_HD Cons<T> physical_flux(Local<T> l)
{
    Cons<T> f;
    f.rho = l.m.x;
    f.m.x = (l.m.x * l.u.x) + l.p;
    f.m.y = (l.m.x * l.u.y);
    f.m.z = (l.m.x * l.u.z);
    f.e = (l.e + l.p) * l.u.x;
    return f;
}

template <typename T> TEAL_HD Cons<T> jump(Local<T> l, Local<T> r) // convective.c:103-110
{
    Cons<T> j;
    j.rho = r.rho - l.rho;
    j.m = sub(r.m, l.m);
    j.e = r.e - l.e;
    return j;
}

template <typename T> TEAL_HD V3<T> roe_velocity(T w, T sl, T sr, Local<T> l, Local<T> r)
{
    return mul(w, add(mul(sl, l.u), mul(sr, r.u)));
}

// ---- exact Riemann solver (src/euler/riemann.c) -----------------------------------
// `err` is set to 1 (vacuum) or 2 (no convergence) where the reference aborts.
template <typename T>
TEAL_HD T guess_pressure(const T *l, const T *r, T lc, T rc, const T *gam)
{
    T cup = (l[0] + r[0]) * (lc + rc) / 4;
    T diff = l[1] - r[1];
    T pvrs = tmax((l[4] + r[4]) + (diff * cup), T(1e-6)) / 2;
    T pmin = tmin(l[4], r[4]);
    T pmax = tmax(l[4], r[4]);
    T qmax = pmax / pmin;
    if (qmax <= 2 && pmin <= pvrs && pvrs <= pmax) {
        return pvrs;
    }
    if (pvrs < pmin) {
        T ratio = pow(l[4] / r[4], gam[1]);
        T guess = ((ratio * l[1] / lc) + (r[1] / rc) + (gam[4] * (ratio - 1))) /
                  ((ratio / lc) + (1 / rc));
        T tl = 1 + (gam[7] * (l[1] - guess) / lc);
        T trr = 1 + (gam[7] * (guess - r[1]) / rc);
        return ((l[4] * pow(tl, gam[3])) + (r[4] * pow(trr, gam[3]))) / 2;
    }
    T sl = tsqrt((gam[5] / l[0]) / ((gam[6] * l[4]) + pvrs));
    T sr = tsqrt((gam[5] / r[0]) / ((gam[6] * r[4]) + pvrs));
    return ((sl * l[4]) + (sr * r[4]) + diff) / (sl + sr);
}

template <typename T>
TEAL_HD void prefun(T *f, T *df, T pold, const T *s, T c, const T *gam)
{
    if (pold <= s[4]) {
        T ratio = pold / s[4];
        *f = gam[4] * c * (pow(ratio, gam[1]) - 1);
        *df = (1 / (s[0] * c)) * pow(ratio, -gam[2]);
    }
    else {
        T a = gam[5] / s[0];
        T b = gam[6] * s[4];
        T root = tsqrt(a / (b + pold));
        *f = (pold - s[4]) * root;
        *df = (1 - ((pold - s[4]) / (b + pold) / 2)) * root;
    }
}

template <typename T>
TEAL_HD void riemann_sample(const T *l, const T *r, T lc, T rc, T su, T sp, const T *gam,
                            T loc, T *o)
{
    o[2] = o[3] = 0;
    if (loc <= su) {
        if (sp <= l[4]) {
            if (loc <= l[1] - lc) {
                for (int i = 0; i < 5; i++) o[i] = l[i];
                return;
            }
            T sc = lc * pow(sp / l[4], gam[1]);
            if (loc > su - sc) {
                o[0] = l[0] * pow(sp / l[4], 1 / gam[0]);
                o[1] = su;
                o[4] = sp;
                return;
            }
            T fc = gam[5] * (lc + (gam[7] * (l[1] - loc)));
            o[0] = l[0] * pow(fc / lc, gam[4]);
            o[1] = gam[5] * (lc + (gam[7] * l[1]) + loc);
            o[4] = l[4] * pow(fc / lc, gam[3]);
            return;
        }
        T ratio = sp / l[4];
        if (loc <= l[1] - (lc * tsqrt((gam[2] * ratio) + gam[1]))) {
            for (int i = 0; i < 5; i++) o[i] = l[i];
            return;
        }
        o[0] = l[0] * (ratio + gam[6]) / ((ratio * gam[6]) + 1);
        o[1] = su;
        o[4] = sp;
        return;
    }
    if (sp > r[4]) {
        T ratio = sp / r[4];
        if (loc >= r[1] + (rc * tsqrt((gam[2] * ratio) + gam[1]))) {
            for (int i = 0; i < 5; i++) o[i] = r[i];
            return;
        }
        o[0] = r[0] * (ratio + gam[6]) / ((ratio * gam[6]) + 1);
        o[1] = su;
        o[4] = sp;
        return;
    }
    if (loc >= r[1] + rc) {
        for (int i = 0; i < 5; i++) o[i] = r[i];
        return;
    }
    T sc = rc * pow(sp / r[4], gam[1]);
    if (loc <= su + sc) {
        o[0] = r[0] * pow(sp / r[4], 1 / gam[0]);
        o[1] = su;
        o[4] = sp;
        return;
    }
    T fc = gam[5] * (rc - (gam[7] * (r[1] - loc)));
    o[0] = r[0] * pow(fc / rc, gam[4]);
    o[1] = gam[5] * (-rc + (gam[7] * r[1]) + loc);
    o[4] = r[4] * pow(fc / rc, gam[3]);
}

template <typename T>
TEAL_HD void riemann(const T *l, const T *r, T gamma, T loc, T *out, int *err)
{
    T gam[9] = {gamma,
                (gamma - 1) / (2 * gamma),
                (gamma + 1) / (2 * gamma),
                2 * gamma / (gamma - 1),
                2 / (gamma - 1),
                2 / (gamma + 1),
                (gamma - 1) / (gamma + 1),
                (gamma - 1) / 2,
                gamma - 1};
    T lc = tsqrt(gamma * l[4] / l[0]);
    T rc = tsqrt(gamma * r[4] / r[0]);
    if (gam[4] * (lc + rc) <= (r[1] - l[1])) {
        *err = 1;
        for (int i = 0; i < 5; i++) out[i] = 0;
        return;
    }
    T diff = r[1] - l[1];
    T pold = guess_pressure(l, r, lc, rc, gam);
    T su = 0;
    T sp = 0;
    int done = 0;
    for (int iter = 0; iter < 20; iter++) {
        T lf, ldf, rf, rdf;
        prefun(&lf, &ldf, pold, l, lc, gam);
        prefun(&rf, &rdf, pold, r, rc, gam);
        T pnew = pold - ((lf + rf + diff) / (ldf + rdf));
        if (2 * tabs((pnew - pold) / (pnew + pold)) <= T(1e-6)) {
            su = (l[1] + r[1] + rf - lf) / 2;
            sp = pnew;
            done = 1;
            break;
        }
        pold = tmax(T(1e-6), pnew);
    }
    if (!done) {
        *err = 2;
    }
    riemann_sample(l, r, lc, rc, su, sp, gam, loc, out);
    if (loc <= su) {
        out[2] = l[2];
        out[3] = l[3];
    }
    else {
        out[2] = r[2];
        out[3] = r[3];
    }
}

// ---- convective fluxes (src/euler/convective.c) -----------------------------------
template <typename T>
TEAL_HD void flux_godunov(T *flux, const T *lp, const T *rp, T gamma, M3<T> b, int *err)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T lprim[5] = {l.rho, l.u.x, l.u.y, l.u.z, l.p};
    T rprim[5] = {r.rho, r.u.x, r.u.y, r.u.z, r.p};
    T fp[5];
    riemann(lprim, rprim, gamma, T(0), fp, err);
    Local<T> f;
    f.rho = fp[0];
    f.u = v3(fp[1], fp[2], fp[3]);
    f.p = fp[4];
    f.m = mul(f.rho, f.u);
    f.e = (f.p / (gamma - 1)) + (dot(f.m, f.u) / 2);
    to_global(flux, physical_flux(f), b);
}

template <typename T> TEAL_HD void entropy_fix(T *eig, Local<T> l, Local<T> r, T gamma)
{
    T cl = tsqrt(gamma * l.p / l.rho);
    T cr = tsqrt(gamma * r.p / r.rho);
    T el[3] = {l.u.x - cl, l.u.x, l.u.x + cl};
    T er[3] = {r.u.x - cr, r.u.x, r.u.x + cr};
    for (int i = 0; i < 3; i++) {
        T lambda = eig[i];
        T delta = tmax(T(0), tmax(lambda - el[i], er[i] - lambda));
        eig[i] = (tabs(lambda) < delta) ? delta : tabs(lambda);
    }
}

template <typename T> TEAL_HD void flux_roe(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T gm1 = gamma - 1;
    T hl = (l.e + l.p) / l.rho;
    T hr = (r.e + r.p) / r.rho;
    T sl = tsqrt(l.rho);
    T sr = tsqrt(r.rho);
    T w = 1 / (sl + sr);
    V3<T> u = roe_velocity(w, sl, sr, l, r);
    T h = w * ((sl * hl) + (sr * hr));
    T u2 = dot(u, u);
    T c2 = gm1 * (h - (u2 / 2));
    T c = tsqrt(c2);
    T eig[3] = {u.x - c, u.x, u.x + c};
    entropy_fix(eig, l, r, gamma);
    T ws[5];
    Cons<T> j = jump(l, r);
    ws[2] = j.m.y - (u.y * j.rho);
    ws[3] = j.m.z - (u.z * j.rho);
    ws[1] = gm1 / c2 *
            ((j.rho * (h - sqr(u.x))) + (u.x * j.m.x) - j.e + (ws[2] * u.y) + (ws[3] * u.z));
    ws[0] = ((j.rho * (u.x + c)) - j.m.x - (c * ws[1])) / (2 * c);
    ws[4] = j.rho - (ws[0] + ws[1]);
    T ev[5][5] = {
        {1, 1, 0, 0, 1},
        {u.x - c, u.x, 0, 0, u.x + c},
        {u.y, u.y, 1, 0, u.y},
        {u.z, u.z, 0, 1, u.z},
        {h - (u.x * c), u2 / 2, u.y, u.z, h + (u.x * c)},
    };
    T ch[5] = {ws[0] * eig[0], ws[1] * eig[1], ws[2] * eig[1], ws[3] * eig[1], ws[4] * eig[2]};
    T dis[5];
    for (int i = 0; i < 5; i++) {
        dis[i] = 0;
        for (int k = 0; k < 5; k++) {
            dis[i] += ev[i][k] * ch[k];
        }
    }
    Cons<T> fl = physical_flux(l);
    Cons<T> fr = physical_flux(r);
    Cons<T> f;
    f.rho = (fl.rho + fr.rho - dis[0]) / 2;
    f.m.x = (fl.m.x + fr.m.x - dis[1]) / 2;
    f.m.y = (fl.m.y + fr.m.y - dis[2]) / 2;
    f.m.z = (fl.m.z + fr.m.z - dis[3]) / 2;
    f.e = (fl.e + fr.e - dis[4]) / 2;
    to_global(flux, f, b);
}

template <typename T> TEAL_HD Cons<T> average_flux(Local<T> l, Local<T> r, T ssl, T ssr)
{
    Cons<T> fl = physical_flux(l);
    fl.rho *= ssr;
    fl.m.x *= ssr;
    fl.m.y *= ssr;
    fl.m.z *= ssr;
    fl.e *= ssr;
    Cons<T> fr = physical_flux(r);
    fr.rho *= ssl;
    fr.m.x *= ssl;
    fr.m.y *= ssl;
    fr.m.z *= ssl;
    fr.e *= ssl;
    T ssj = ssl * ssr;
    Cons<T> j = jump(l, r);
    j.rho *= ssj;
    j.m.x *= ssj;
    j.m.y *= ssj;
    j.m.z *= ssj;
    j.e *= ssj;
    T factor = 1 / (ssr - ssl);
    Cons<T> f;
    f.rho = factor * (fl.rho - fr.rho + j.rho);
    f.m = mul(factor, add(sub(fl.m, fr.m), j.m));
    f.e = factor * (fl.e - fr.e + j.e);
    return f;
}

template <typename T>
TEAL_HD void roe_signal_speeds(Local<T> l, Local<T> r, T gamma, T *ssl, T *ssr)
{
    T hl = (l.e + l.p) / l.rho;
    T hr = (r.e + r.p) / r.rho;
    T sl = tsqrt(l.rho);
    T sr = tsqrt(r.rho);
    T w = 1 / (sl + sr);
    V3<T> u = roe_velocity(w, sl, sr, l, r);
    T h = w * ((sl * hl) + (sr * hr));
    T u2 = dot(u, u);
    T c = tsqrt((gamma - 1) * (h - (u2 / 2)));
    T cl = tsqrt(gamma * l.p / l.rho);
    T cr = tsqrt(gamma * r.p / r.rho);
    *ssl = tmin(l.u.x - cl, u.x - c);
    *ssr = tmax(r.u.x + cr, u.x + c);
}

template <typename T> TEAL_HD void flux_hll(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T ssl, ssr;
    roe_signal_speeds(l, r, gamma, &ssl, &ssr);
    Cons<T> f;
    if (ssl >= 0) {
        f = physical_flux(l);
    }
    else if (ssr <= 0) {
        f = physical_flux(r);
    }
    else {
        f = average_flux(l, r, ssl, ssr);
    }
    to_global(flux, f, b);
}

template <typename T> TEAL_HD Cons<T> contact_flux(Local<T> k, T ssk, T fk, T ss)
{
    T factor = fk / (ssk - ss);
    Cons<T> star;
    star.rho = factor;
    star.m.x = factor * ss;
    star.m.y = factor * k.u.y;
    star.m.z = factor * k.u.z;
    star.e = factor * ((k.e / k.rho) + ((ss - k.u.x) * (ss + (k.p / fk))));
    Cons<T> f = physical_flux(k);
    f.rho += ssk * (star.rho - k.rho);
    V3<T> dm = mul(ssk, sub(star.m, k.m));
    f.m.x += dm.x;
    f.m.y += dm.y;
    f.m.z += dm.z;
    f.e += ssk * (star.e - k.e);
    return f;
}

template <typename T> TEAL_HD void flux_hllc(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T ssl, ssr;
    roe_signal_speeds(l, r, gamma, &ssl, &ssr);
    // the branches of convective.c:279-346 pick a side k and a formula; the side is selected
    // first so that lanes of one wave taking different sides evaluate one formula, not both
    Cons<T> f;
    const bool super = (ssl >= 0) || (ssr <= 0);
    T ss = 0, ssk = 0, fk = 0;
    bool left = ssl >= 0;
    if (!super) {
        T fl = l.rho * (ssl - l.u.x);
        T fr = r.rho * (ssr - r.u.x);
        ss = (r.p - l.p + (fl * l.u.x) - (fr * r.u.x)) / (fl - fr);
        left = (ssl <= 0 && 0 <= ss);
        ssk = left ? ssl : ssr;
        fk = left ? fl : fr;
    }
    Local<T> k;
    k.rho = left ? l.rho : r.rho;
    k.u = v3(left ? l.u.x : r.u.x, left ? l.u.y : r.u.y, left ? l.u.z : r.u.z);
    k.p = left ? l.p : r.p;
    k.m = v3(left ? l.m.x : r.m.x, left ? l.m.y : r.m.y, left ? l.m.z : r.m.z);
    k.e = left ? l.e : r.e;
    if (super) {
        f = physical_flux(k);
    }
    else {
        f = contact_flux(k, ssk, fk, ss);
    }
    to_global(flux, f, b);
}

// flux_hllc without the face frame, for the FP32 path: HLLC is invariant under rotations about
// the normal n, so the flux rotated back to the global frame needs n alone.  With u_n = u.n,
// the physical flux of a state is {rho u_n, rho u_n u + p n, (e + p) u_n} and the contact
// state's momentum is rho* (u + (s* - u_n) n) (tangential velocity kept, normal velocity s*);
// the signal speeds take the normal components of the same Roe averages.  The FP64 path keeps
// flux_hllc, whose face-frame arithmetic is the reference's (convective.c:279-346).
template <typename T>
struct Cart {
    T rho, un, p, e;
    V3<T> u, m;
};

template <typename T> TEAL_HD Cart<T> to_cart(const T *g, T gamma, V3<T> n)
{
    Cart<T> s;
    s.rho = g[0];
    s.u = v3(g[1], g[2], g[3]);
    s.p = g[4];
    s.un = dot(s.u, n);
    s.m = mul(s.rho, s.u);
    s.e = (s.p / (gamma - 1)) + (dot(s.m, s.u) / 2);
    return s;
}

template <typename T> TEAL_HD void flux_hllc_normal(T *flux, const T *lp, const T *rp, T gamma, V3<T> n)
{
    const Cart<T> l = to_cart(lp, gamma, n);
    const Cart<T> r = to_cart(rp, gamma, n);
    T hl = (l.e + l.p) / l.rho;
    T hr = (r.e + r.p) / r.rho;
    T sl = tsqrt(l.rho);
    T sr = tsqrt(r.rho);
    T w = 1 / (sl + sr);
    V3<T> u = mul(w, add(mul(sl, l.u), mul(sr, r.u)));
    T h = w * ((sl * hl) + (sr * hr));
    T c = tsqrt((gamma - 1) * (h - (dot(u, u) / 2)));
    T cl = tsqrt(gamma * l.p / l.rho);
    T cr = tsqrt(gamma * r.p / r.rho);
    T un = dot(u, n);
    T ssl = tmin(l.un - cl, un - c);
    T ssr = tmax(r.un + cr, un + c);
    const bool super = (ssl >= 0) || (ssr <= 0);
    T ss = 0, ssk = 0, fk = 0;
    bool left = ssl >= 0;
    if (!super) {
        T fl = l.rho * (ssl - l.un);
        T fr = r.rho * (ssr - r.un);
        ss = (r.p - l.p + (fl * l.un) - (fr * r.un)) / (fl - fr);
        left = (ssl <= 0 && 0 <= ss);
        ssk = left ? ssl : ssr;
        fk = left ? fl : fr;
    }
    const Cart<T> k = left ? l : r;
    T mn = k.rho * k.un;
    T frho = mn;
    V3<T> fm = add(mul(mn, k.u), mul(k.p, n));
    T fe = (k.e + k.p) * k.un;
    if (!super) {
        T factor = fk / (ssk - ss);
        V3<T> mstar = mul(factor, add(k.u, mul(ss - k.un, n)));
        T estar = factor * ((k.e / k.rho) + ((ss - k.un) * (ss + (k.p / fk))));
        frho += ssk * (factor - k.rho);
        fm = add(fm, mul(ssk, sub(mstar, k.m)));
        fe += ssk * (estar - k.e);
    }
    flux[0] = frho;
    flux[1] = fm.x;
    flux[2] = fm.y;
    flux[3] = fm.z;
    flux[4] = fe;
}

template <typename T> TEAL_HD void flux_hlle(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T sl = tsqrt(l.rho);
    T sr = tsqrt(r.rho);
    T w = 1 / (sl + sr);
    T ux = w * ((sl * l.u.x) + (sr * r.u.x));
    T c2l = gamma * l.p / l.rho;
    T c2r = gamma * r.p / r.rho;
    T cl = tsqrt(c2l);
    T cr = tsqrt(c2r);
    T c2 = w * ((sl * c2l) + (sr * c2r));
    T vj = sqr(w) * sl * sr * sqr(r.u.x - l.u.x) / 2;
    T c = tsqrt(c2 + vj);
    T ssl = tmin(l.u.x - cl, ux - c);
    T ssr = tmax(r.u.x + cr, ux + c);
    Cons<T> f;
    if (ssl >= 0) {
        f = physical_flux(l);
    }
    else if (ssr <= 0) {
        f = physical_flux(r);
    }
    else {
        f = average_flux(l, r, ssl, ssr);
    }
    to_global(flux, f, b);
}

// split Mach/pressure functions of ausmd/ausmdv (convective.c:399-438, 471-510);
// VFLAG selects ausmdv's form of the supersonic pressure term
template <typename T, int VFLAG>
TEAL_HD void ausm_split(Local<T> l, Local<T> r, T gamma, T *vpl, T *ppl, T *vmr, T *pmr)
{
    T cl = tsqrt(gamma * l.p / l.rho);
    T cr = tsqrt(gamma * r.p / r.rho);
    T c = tmax(cl, cr);
    T pdl = l.p / l.rho;
    T pdr = r.p / r.rho;
    T pds = pdl + pdr;
    T ml = l.u.x / c;
    T mr = r.u.x / c;
    if (tabs(l.u.x) <= c) {
        T al = 2 * pdl / pds;
        T fac = l.u.x + c;
        *vpl = (al * sqr(fac) / (4 * c)) + ((1 - al) * (l.u.x + tabs(l.u.x)) / 2);
        *ppl = (l.p * sqr(fac) / (4 * sqr(c))) * (2 - ml);
    }
    else {
        *vpl = (l.u.x + tabs(l.u.x)) / 2;
        *ppl = VFLAG ? (l.p * (l.u.x + tabs(l.u.x)) / (2 * l.u.x))
                     : ((l.p * (l.u.x + tabs(l.u.x)) / 2) / l.u.x);
    }
    if (tabs(r.u.x) <= c) {
        T ar = 2 * pdr / pds;
        T fac = r.u.x - c;
        *vmr = (-ar * sqr(fac) / (4 * c)) + ((1 - ar) * (r.u.x - tabs(r.u.x)) / 2);
        *pmr = (r.p * sqr(fac) / (4 * sqr(c))) * (2 + mr);
    }
    else {
        *vmr = (r.u.x - tabs(r.u.x)) / 2;
        *pmr = VFLAG ? (r.p * (r.u.x - tabs(r.u.x)) / (2 * r.u.x))
                     : ((r.p * (r.u.x - tabs(r.u.x)) / 2) / r.u.x);
    }
}

template <typename T> TEAL_HD void flux_ausmd(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T vpl, ppl, vmr, pmr;
    ausm_split<T, 0>(l, r, gamma, &vpl, &ppl, &vmr, &pmr);
    T mf = (vpl * l.rho) + (vmr * r.rho);
    T amf = tabs(mf);
    T pf = ppl + pmr;
    T hl = (l.e + l.p) / l.rho;
    T hr = (r.e + r.p) / r.rho;
    Cons<T> f;
    f.rho = mf;
    f.m.x = (((mf * (l.u.x + r.u.x)) - (amf * (r.u.x - l.u.x))) / 2) + pf;
    f.m.y = ((mf * (l.u.y + r.u.y)) - (amf * (r.u.y - l.u.y))) / 2;
    f.m.z = ((mf * (l.u.z + r.u.z)) - (amf * (r.u.z - l.u.z))) / 2;
    f.e = ((mf * (hl + hr)) - (amf * (hr - hl))) / 2;
    to_global(flux, f, b);
}

template <typename T> TEAL_HD void flux_ausmdv(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T vpl, ppl, vmr, pmr;
    ausm_split<T, 1>(l, r, gamma, &vpl, &ppl, &vmr, &pmr);
    T mf = (vpl * l.rho) + (vmr * r.rho);
    T amf = tabs(mf);
    T pf = ppl + pmr;
    T nd = ((mf * (l.u.x + r.u.x)) - (amf * (r.u.x - l.u.x))) / 2;
    T nv = (vpl * l.m.x) + (vmr * r.m.x);
    T pmin = tmin(l.p, r.p);
    T sw = tmin(T(1.0) / 2, tmin(T(1), (T(10) * tabs(r.p - l.p)) / pmin));
    T nm = (((1 + sw) * nv) + ((1 - sw) * nd)) / 2;
    T hl = (l.e + l.p) / l.rho;
    T hr = (r.e + r.p) / r.rho;
    Cons<T> f;
    f.rho = mf;
    f.m.x = nm + pf;
    f.m.y = ((mf * (l.u.y + r.u.y)) - (amf * (r.u.y - l.u.y))) / 2;
    f.m.z = ((mf * (l.u.z + r.u.z)) - (amf * (r.u.z - l.u.z))) / 2;
    f.e = ((mf * (hl + hr)) - (amf * (hr - hl))) / 2;
    to_global(flux, f, b);
}

template <typename T> TEAL_HD void flux_lxf(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    T cl = tsqrt(gamma * l.p / l.rho);
    T cr = tsqrt(gamma * r.p / r.rho);
    T ss = tmax(tabs(l.u.x) + cl, tabs(r.u.x) + cr);
    Cons<T> j = jump(l, r);
    j.rho *= ss;
    j.m.x *= ss;
    j.m.y *= ss;
    j.m.z *= ss;
    j.e *= ss;
    Cons<T> fl = physical_flux(l);
    Cons<T> fr = physical_flux(r);
    Cons<T> f;
    f.rho = (fl.rho + fr.rho - j.rho) / 2;
    f.m = divv(sub(add(fl.m, fr.m), j.m), T(2));
    f.e = (fl.e + fr.e - j.e) / 2;
    to_global(flux, f, b);
}

template <typename T> TEAL_HD void flux_central(T *flux, const T *lp, const T *rp, T gamma, M3<T> b)
{
    Local<T> l = to_local(lp, gamma, b);
    Local<T> r = to_local(rp, gamma, b);
    Cons<T> fl = physical_flux(l);
    Cons<T> fr = physical_flux(r);
    Cons<T> f;
    f.rho = (fl.rho + fr.rho) / 2;
    f.m = divv(add(fl.m, fr.m), T(2));
    f.e = (fl.e + fr.e) / 2;
    to_global(flux, f, b);
}

// compile-time flux dispatch (one kernel instantiation per flux)
template <int FLUX, typename T>
TEAL_HD void convective(T *flux, const T *l, const T *r, T gamma, M3<T> b, int *err)
{
    if constexpr (FLUX == FLUX_GODUNOV) flux_godunov(flux, l, r, gamma, b, err);
    else if constexpr (FLUX == FLUX_ROE) flux_roe(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_HLL) flux_hll(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_HLLC) flux_hllc(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_HLLE) flux_hlle(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_AUSMD) flux_ausmd(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_AUSMDV) flux_ausmdv(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_LXF) flux_lxf(flux, l, r, gamma, b);
    else if constexpr (FLUX == FLUX_CENTRAL) flux_central(flux, l, r, gamma, b);
    else {
        for (int i = 0; i < 5; i++) flux[i] = 0;
    }
}

// runtime dispatch for host callers and tests
template <typename T>
TEAL_HD void convective_rt(int kind, T *flux, const T *l, const T *r, T gamma, M3<T> b, int *err)
{
    switch (kind) {
        case FLUX_GODUNOV: convective<FLUX_GODUNOV>(flux, l, r, gamma, b, err); return;
        case FLUX_ROE: convective<FLUX_ROE>(flux, l, r, gamma, b, err); return;
        case FLUX_HLL: convective<FLUX_HLL>(flux, l, r, gamma, b, err); return;
        case FLUX_HLLC: convective<FLUX_HLLC>(flux, l, r, gamma, b, err); return;
        case FLUX_HLLE: convective<FLUX_HLLE>(flux, l, r, gamma, b, err); return;
        case FLUX_AUSMD: convective<FLUX_AUSMD>(flux, l, r, gamma, b, err); return;
        case FLUX_AUSMDV: convective<FLUX_AUSMDV>(flux, l, r, gamma, b, err); return;
        case FLUX_LXF: convective<FLUX_LXF>(flux, l, r, gamma, b, err); return;
        case FLUX_CENTRAL: convective<FLUX_CENTRAL>(flux, l, r, gamma, b, err); return;
        default: convective<FLUX_NONE>(flux, l, r, gamma, b, err); return;
    }
}

// ---- viscous central flux (src/navierstokes/viscous.c:11-39) ------------------------
// g = {grad rho, grad u, grad v, grad w, grad p} (15 values)
template <typename T>
TEAL_HD void viscous_central(T *flux, const T *p, const V3<T> *g, T gamma, T mu, T pr, V3<T> n)
{
    M3<T> tau;
    T div = g[1].x + g[2].y + g[3].z;
    tau.x.x = 2 * mu * (g[1].x - (div / 3));
    tau.y.y = 2 * mu * (g[2].y - (div / 3));
    tau.z.z = 2 * mu * (g[3].z - (div / 3));
    tau.y.x = tau.x.y = mu * (g[1].y + g[2].x);
    tau.z.x = tau.x.z = mu * (g[1].z + g[3].x);
    tau.z.y = tau.y.z = mu * (g[2].z + g[3].y);
    T factor = -mu * gamma / ((gamma - 1) * pr * sqr(p[0]));
    V3<T> heat = mul(factor, sub(mul(p[0], g[4]), mul(p[4], g[0])));
    V3<T> theta = sub(mv(tau, v3(p[1], p[2], p[3])), heat);
    V3<T> m = mv(tau, n);
    flux[0] = 0;
    flux[1] = m.x;
    flux[2] = m.y;
    flux[3] = m.z;
    flux[4] = dot(theta, n);
}

// viscous_central for the FP32 path, from the same tau with the products the flux needs
// taken along n first: tau is symmetric, so the energy flux (tau u - q).n = (tau n).u - q.n,
// and the heat flux enters only as q.n = factor (rho (grad p).n - p (grad rho).n).  The FP64
// path keeps viscous_central, the reference's arithmetic.
template <typename T>
TEAL_HD void viscous_central_normal(T *flux, const T *p, const V3<T> *g, T gamma, T mu, T pr, V3<T> n)
{
    M3<T> tau;
    T div = g[1].x + g[2].y + g[3].z;
    tau.x.x = 2 * mu * (g[1].x - (div / 3));
    tau.y.y = 2 * mu * (g[2].y - (div / 3));
    tau.z.z = 2 * mu * (g[3].z - (div / 3));
    tau.y.x = tau.x.y = mu * (g[1].y + g[2].x);
    tau.z.x = tau.x.z = mu * (g[1].z + g[3].x);
    tau.z.y = tau.y.z = mu * (g[2].z + g[3].y);
    T factor = -mu * gamma / ((gamma - 1) * pr * sqr(p[0]));
    T qn = factor * ((p[0] * dot(g[4], n)) - (p[4] * dot(g[0], n)));
    V3<T> m = mv(tau, n);
    flux[0] = 0;
    flux[1] = m.x;
    flux[2] = m.y;
    flux[3] = m.z;
    flux[4] = dot(m, v3(p[1], p[2], p[3])) - qn;
}

// ---- boundary conditions (src/euler/boundary.c:7-271, navierstokes/boundary.c:6-18) --
template <typename T> TEAL_HD void bc_to_global(T *g, Local<T> l, T gamma, M3<T> b)
{
    l.u = divv(l.m, l.rho);
    l.p = (gamma - 1) * (l.e - (dot(l.m, l.u) / 2));
    g[0] = tmax(T(1e-8), l.rho);
    V3<T> v = mv(tr(b), l.u);
    g[1] = v.x;
    g[2] = v.y;
    g[3] = v.z;
    g[4] = tmax(T(1e-8), l.p);
}

template <typename T>
TEAL_HD void bc_farfield(T *outer, const T *inner_, T gamma, M3<T> b, const T *ctx)
{
    Local<T> ol = to_local(ctx, gamma, b);
    Local<T> in = to_local(inner_, gamma, b);
    T gm1 = gamma - 1;
    T v2 = dot(in.u, in.u);
    T c2 = gamma * in.p / in.rho;
    T c = tsqrt(c2);
    T h = (v2 / 2) + (c2 / gm1);
    V3<T> u = in.u;
    T evi[5][5] = {
        {h + (c / gm1 * (u.x - c)), -(u.x + (c / gm1)), -u.y, -u.z, 1},
        {(-2 * h) + (4 / gm1 * c2), 2 * u.x, 2 * u.y, 2 * u.z, -2},
        {-2 * u.y * c2 / gm1, 0, 2 * c2 / gm1, 0, 0},
        {-2 * u.z * c2 / gm1, 0, 0, 2 * c2 / gm1, 0},
        {h - (c / gm1 * (u.x + c)), -u.x + (c / gm1), -u.y, -u.z, 1},
    };
    T factor = gm1 / (2 * c2);
    T ui[5] = {in.rho, in.m.x, in.m.y, in.m.z, in.e};
    T ug[5] = {ol.rho, ol.m.x, ol.m.y, ol.m.z, ol.e};
    T chi[5], chg[5];
    for (int i = 0; i < 5; i++) {
        chi[i] = 0;
        chg[i] = 0;
        for (int j = 0; j < 5; j++) {
            chi[i] += evi[i][j] * ui[j];
            chg[i] += evi[i][j] * ug[j];
        }
        chi[i] *= factor;
        chg[i] *= factor;
    }
    T cg = tsqrt(gamma * ol.p / ol.rho);
    if (ol.u.x - cg > 0) {
        chg[0] = chi[0];
    }
    if (ol.u.x > 0) {
        chg[1] = chi[1];
        chg[2] = chi[2];
        chg[3] = chi[3];
    }
    if (ol.u.x + cg > 0) {
        chg[4] = chi[4];
    }
    T ev[5][5] = {
        {1, 1, 0, 0, 1},
        {u.x - c, u.x, 0, 0, u.x + c},
        {u.y, u.y, 1, 0, u.y},
        {u.z, u.z, 0, 1, u.z},
        {h - (u.x * c), v2 / 2, u.y, u.z, h + (u.x * c)},
    };
    for (int i = 0; i < 5; i++) {
        ug[i] = 0;
        for (int j = 0; j < 5; j++) {
            ug[i] += ev[i][j] * chg[j];
        }
    }
    ol.rho = ug[0];
    ol.m = v3(ug[1], ug[2], ug[3]);
    ol.e = ug[4];
    bc_to_global(outer, ol, gamma, b);
}

template <typename T>
TEAL_HD void boundary(int kind, T *outer, const T *inner, T gamma, M3<T> b, const T *ctx)
{
    V3<T> vi = v3(inner[1], inner[2], inner[3]);
    switch (kind) {
        case BC_SYMMETRY: {
            T vel = dot(vi, b.x);
            V3<T> v = sub(vi, mul(2 * vel, b.x));
            outer[0] = inner[0];
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            outer[4] = inner[4];
            return;
        }
        case BC_SUPERSONIC_INFLOW:
            for (int i = 0; i < 5; i++) outer[i] = ctx[i];
            return;
        case BC_SUPERSONIC_OUTFLOW:
            for (int i = 0; i < 5; i++) outer[i] = inner[i];
            return;
        case BC_SUBSONIC_INFLOW: {
            T density = inner[0];
            T c2 = gamma * inner[4] / inner[0];
            T c = tsqrt(c2);
            V3<T> vc = v3(ctx[1], ctx[2], ctx[3]);
            T vel = dot(sub(vc, vi), b.x);
            T f1 = density * c;
            outer[4] = (ctx[4] + inner[4] - (f1 * vel)) / 2;
            outer[0] = ctx[0] + ((outer[4] - ctx[4]) / c2);
            T f2 = (ctx[4] - outer[4]) / f1;
            V3<T> v = sub(vc, mul(f2, b.x));
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            return;
        }
        case BC_SUBSONIC_OUTFLOW: {
            T density = inner[0];
            T c2 = gamma * inner[4] / inner[0];
            T c = tsqrt(c2);
            T pressure = ctx[4];
            outer[4] = pressure;
            outer[0] = inner[0] + ((pressure - inner[4]) / c2);
            T f = (inner[4] - pressure) / (density * c);
            V3<T> v = sub(vi, mul(f, b.x));
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            return;
        }
        case BC_PRESSURE_OUTFLOW: {
            T c = tsqrt(gamma * inner[4] / inner[0]);
            T vel = dot(vi, b.x);
            T pressure = (vel / c < 1) ? ctx[4] : inner[4];
            outer[0] = inner[0] * pressure / inner[4];
            outer[1] = inner[1];
            outer[2] = inner[2];
            outer[3] = inner[3];
            outer[4] = pressure;
            return;
        }
        case BC_FARFIELD: bc_farfield(outer, inner, gamma, b, ctx); return;
        case BC_ADIABATIC_WALL: {
            V3<T> v = mul(T(-1), vi);
            outer[0] = inner[0];
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            outer[4] = inner[4];
            return;
        }
        default:
            for (int i = 0; i < 5; i++) outer[i] = inner[i];
            return;
    }
}

// ---- conversions and time steps (src/euler/create.c:6-44, navierstokes/create.c:7-26)
template <typename T> TEAL_HD void cons2prim(T *p, const T *u, T gamma)
{
    p[0] = tmax(T(1e-8), u[0]);
    V3<T> v = divv(v3(u[1], u[2], u[3]), p[0]);
    p[1] = v.x;
    p[2] = v.y;
    p[3] = v.z;
    p[4] = tmax(T(1e-8), (gamma - 1) * (u[4] - (dot(v3(u[1], u[2], u[3]), v) / 2)));
}

// cons2prim with one reciprocal of the density in place of three divisions (the FP32 path's
// FP64 update, which is not held to the reference's rounding)
template <typename T> TEAL_HD void cons2prim_recip(T *p, const T *u, T gamma)
{
    p[0] = tmax(T(1e-8), u[0]);
    const T ir = T(1) / p[0];
    V3<T> v = v3(u[1] * ir, u[2] * ir, u[3] * ir);
    p[1] = v.x;
    p[2] = v.y;
    p[3] = v.z;
    p[4] = tmax(T(1e-8), (gamma - 1) * (u[4] - (dot(v3(u[1], u[2], u[3]), v) / 2)));
}

template <typename T> TEAL_HD void prim2cons(T *u, const T *p, T gamma)
{
    u[0] = p[0];
    V3<T> m = mul(u[0], v3(p[1], p[2], p[3]));
    u[1] = m.x;
    u[2] = m.y;
    u[3] = m.z;
    u[4] = (p[4] / (gamma - 1)) + (dot(m, v3(p[1], p[2], p[3])) / 2);
}

template <typename T> TEAL_HD T euler_dt(const T *p, T gamma, T volume, V3<T> proj)
{
    T c = tsqrt(gamma * p[4] / p[0]);
    T lx = (tabs(p[1]) + c) * proj.x;
    T ly = (tabs(p[2]) + c) * proj.y;
    T lz = (tabs(p[3]) + c) * proj.z;
    return volume / (lx + ly + lz);
}

template <typename T> TEAL_HD T ns_dt(const T *p, T gamma, T mu, T pr, T volume, V3<T> proj)
{
    T c = tsqrt(gamma * p[4] / p[0]);
    T lx = (tabs(p[1]) + c) * proj.x;
    T ly = (tabs(p[2]) + c) * proj.y;
    T lz = (tabs(p[3]) + c) * proj.z;
    T lv = tmax(4 * mu / (3 * p[0]), gamma * mu / (p[0] * pr));
    return volume / (lx + ly + lz + (4 * lv * dot(proj, proj) / volume));
}

}  // namespace fv
