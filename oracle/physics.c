/* CPU ORACLE (test infrastructure only) -- Euler / Navier-Stokes plugins.
 *
 * Restates src/euler/convective.c, src/euler/boundary.c, src/euler/create.c,
 * src/euler/riemann.c, src/navierstokes/viscous.c, src/navierstokes/boundary.c and
 * src/navierstokes/create.c.  State layout: primitive {rho, u, v, w, p},
 * conserved {rho, rho u, rho v, rho w, E}.  Association follows the reference. */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define SQ(v) ((v) * (v))

typedef struct {
    double density;
    OVec velocity;
    double pressure;
    OVec momentum;
    double energy;
} Local;

typedef struct {
    double density;
    OVec momentum;
    double energy;
} Cons;

/* src/euler/convective.c:15-25 */
static Local to_local(const double *g, const double *property, OMat basis)
{
    Local l;
    l.density = g[0];
    l.velocity = om_vec(basis, (OVec){g[1], g[2], g[3]});
    l.pressure = g[4];
    l.momentum = ov_mul(l.density, l.velocity);
    l.energy = (l.pressure / (property[0] - 1)) + (ov_dot(l.momentum, l.velocity) / 2);
    return l;
}

/* src/euler/convective.c:27-30 */
static void to_global(double *flux, Cons f, OMat basis)
{
    OVec m = om_vec(om_transpose(basis), f.momentum);
    flux[0] = f.density;
    flux[1] = m.x;
    flux[2] = m.y;
    flux[3] = m.z;
    flux[4] = f.energy;
}

/* src/euler/convective.c:32-41 */
static Cons physical_flux(Local l)
{
    Cons f;
    f.density = l.momentum.x;
    f.momentum.x = (l.momentum.x * l.velocity.x) + l.pressure;
    f.momentum.y = (l.momentum.x * l.velocity.y);
    f.momentum.z = (l.momentum.x * l.velocity.z);
    f.energy = (l.energy + l.pressure) * l.velocity.x;
    return f;
}

static Cons jump_of(Local l, Local r) /* src/euler/convective.c:103-110 */
{
    Cons j;
    j.density = r.density - l.density;
    j.momentum = ov_sub(r.momentum, l.momentum);
    j.energy = r.energy - l.energy;
    return j;
}

static OVec roe_velocity(double w, double sl, double sr, Local l, Local r)
{
    return ov_mul(w, ov_add(ov_mul(sl, l.velocity), ov_mul(sr, r.velocity)));
}

static void godunov(double *flux, const double *lp, const double *rp, const double *property,
                    OMat basis) /* src/euler/convective.c:43-64 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double gamma = property[0];
    double lprim[5] = {l.density, l.velocity.x, l.velocity.y, l.velocity.z, l.pressure};
    double rprim[5] = {r.density, r.velocity.x, r.velocity.y, r.velocity.z, r.pressure};
    double fp[5];
    o_riemann(lprim, rprim, gamma, 0, fp);
    Local face;
    face.density = fp[0];
    face.velocity = (OVec){fp[1], fp[2], fp[3]};
    face.pressure = fp[4];
    face.momentum = ov_mul(face.density, face.velocity);
    face.energy = (face.pressure / (gamma - 1)) + (ov_dot(face.momentum, face.velocity) / 2);
    to_global(flux, physical_flux(face), basis);
}

/* src/euler/convective.c:73-101 */
static void entropy_fix(double *eig, Local l, Local r, double gamma)
{
    double cl = sqrt(gamma * l.pressure / l.density);
    double cr = sqrt(gamma * r.pressure / r.density);
    double el[3] = {l.velocity.x - cl, l.velocity.x, l.velocity.x + cl};
    double er[3] = {r.velocity.x - cr, r.velocity.x, r.velocity.x + cr};
    for (int i = 0; i < 3; i++) {
        double lambda = eig[i];
        double delta = fmax(0, fmax(lambda - el[i], er[i] - lambda));
        if (fabs(lambda) < delta) {
            eig[i] = delta;
        }
        else {
            eig[i] = fabs(lambda);
        }
    }
}

static void roe(double *flux, const double *lp, const double *rp, const double *property,
                OMat basis) /* src/euler/convective.c:112-213 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double gamma = property[0];
    double gm1 = gamma - 1;
    double hl = (l.energy + l.pressure) / l.density;
    double hr = (r.energy + r.pressure) / r.density;
    double sl = sqrt(l.density);
    double sr = sqrt(r.density);
    double w = 1 / (sl + sr);
    OVec u = roe_velocity(w, sl, sr, l, r);
    double h = w * ((sl * hl) + (sr * hr));
    double u2 = ov_norm2(u);
    double c2 = gm1 * (h - (u2 / 2));
    double c = sqrt(c2);
    double eig[3] = {u.x - c, u.x, u.x + c};
    entropy_fix(eig, l, r, gamma);

    double ws[5];
    Cons j = jump_of(l, r);
    ws[2] = j.momentum.y - (u.y * j.density);
    ws[3] = j.momentum.z - (u.z * j.density);
    ws[1] = gm1 / c2 *
            ((j.density * (h - SQ(u.x))) + (u.x * j.momentum.x) - j.energy + (ws[2] * u.y) +
             (ws[3] * u.z));
    ws[0] = ((j.density * (u.x + c)) - j.momentum.x - (c * ws[1])) / (2 * c);
    ws[4] = j.density - (ws[0] + ws[1]);

    double ev[5][5] = {
        {1, 1, 0, 0, 1},
        {u.x - c, u.x, 0, 0, u.x + c},
        {u.y, u.y, 1, 0, u.y},
        {u.z, u.z, 0, 1, u.z},
        {h - (u.x * c), u2 / 2, u.y, u.z, h + (u.x * c)},
    };
    double ch[5] = {ws[0] * eig[0], ws[1] * eig[1], ws[2] * eig[1], ws[3] * eig[1],
                    ws[4] * eig[2]};
    double dis[5];
    for (int i = 0; i < 5; i++) {
        dis[i] = 0;
        for (int k = 0; k < 5; k++) {
            dis[i] += ev[i][k] * ch[k];
        }
    }
    Cons fl = physical_flux(l);
    Cons fr = physical_flux(r);
    Cons f;
    f.density = (fl.density + fr.density - dis[0]) / 2;
    f.momentum.x = (fl.momentum.x + fr.momentum.x - dis[1]) / 2;
    f.momentum.y = (fl.momentum.y + fr.momentum.y - dis[2]) / 2;
    f.momentum.z = (fl.momentum.z + fr.momentum.z - dis[3]) / 2;
    f.energy = (fl.energy + fr.energy - dis[4]) / 2;
    to_global(flux, f, basis);
}

/* src/euler/convective.c:215-239 */
static Cons average_flux(Local l, Local r, double ssl, double ssr)
{
    Cons fl = physical_flux(l);
    fl.density *= ssr;
    fl.momentum.x *= ssr;
    fl.momentum.y *= ssr;
    fl.momentum.z *= ssr;
    fl.energy *= ssr;
    Cons fr = physical_flux(r);
    fr.density *= ssl;
    fr.momentum.x *= ssl;
    fr.momentum.y *= ssl;
    fr.momentum.z *= ssl;
    fr.energy *= ssl;
    double ssj = ssl * ssr;
    Cons j = jump_of(l, r);
    j.density *= ssj;
    j.momentum.x *= ssj;
    j.momentum.y *= ssj;
    j.momentum.z *= ssj;
    j.energy *= ssj;
    double factor = 1 / (ssr - ssl);
    Cons f;
    f.density = factor * (fl.density - fr.density + j.density);
    f.momentum = ov_mul(factor, ov_add(ov_sub(fl.momentum, fr.momentum), j.momentum));
    f.energy = factor * (fl.energy - fr.energy + j.energy);
    return f;
}

/* shared wave-speed estimate of hll/hllc (src/euler/convective.c:249-265, 308-324) */
static void roe_signal_speeds(Local l, Local r, double gamma, double *ssl, double *ssr)
{
    double hl = (l.energy + l.pressure) / l.density;
    double hr = (r.energy + r.pressure) / r.density;
    double sl = sqrt(l.density);
    double sr = sqrt(r.density);
    double w = 1 / (sl + sr);
    OVec u = roe_velocity(w, sl, sr, l, r);
    double h = w * ((sl * hl) + (sr * hr));
    double u2 = ov_norm2(u);
    double c = sqrt((gamma - 1) * (h - (u2 / 2)));
    double cl = sqrt(gamma * l.pressure / l.density);
    double cr = sqrt(gamma * r.pressure / r.density);
    *ssl = fmin(l.velocity.x - cl, u.x - c);
    *ssr = fmax(r.velocity.x + cr, u.x + c);
}

static void hll(double *flux, const double *lp, const double *rp, const double *property,
                OMat basis) /* src/euler/convective.c:241-277 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double ssl;
    double ssr;
    roe_signal_speeds(l, r, property[0], &ssl, &ssr);
    Cons f;
    if (ssl >= 0) {
        f = physical_flux(l);
    }
    else if (ssr <= 0) {
        f = physical_flux(r);
    }
    else {
        f = average_flux(l, r, ssl, ssr);
    }
    to_global(flux, f, basis);
}

/* src/euler/convective.c:279-298 */
static Cons contact_flux(Local k, double ssk, double fk, double ss)
{
    double factor = fk / (ssk - ss);
    Cons star;
    star.density = factor;
    star.momentum.x = factor * ss;
    star.momentum.y = factor * k.velocity.y;
    star.momentum.z = factor * k.velocity.z;
    star.energy =
        factor * ((k.energy / k.density) + ((ss - k.velocity.x) * (ss + (k.pressure / fk))));
    Cons f = physical_flux(k);
    f.density += ssk * (star.density - k.density);
    OVec dm = ov_mul(ssk, ov_sub(star.momentum, k.momentum));
    f.momentum.x += dm.x;
    f.momentum.y += dm.y;
    f.momentum.z += dm.z;
    f.energy += ssk * (star.energy - k.energy);
    return f;
}

static void hllc(double *flux, const double *lp, const double *rp, const double *property,
                 OMat basis) /* src/euler/convective.c:300-346 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double ssl;
    double ssr;
    roe_signal_speeds(l, r, property[0], &ssl, &ssr);
    Cons f;
    if (ssl >= 0) {
        f = physical_flux(l);
    }
    else if (ssr <= 0) {
        f = physical_flux(r);
    }
    else {
        double fl = l.density * (ssl - l.velocity.x);
        double fr = r.density * (ssr - r.velocity.x);
        double ss = (r.pressure - l.pressure + (fl * l.velocity.x) - (fr * r.velocity.x)) /
                    (fl - fr);
        if (ssl <= 0 && 0 <= ss) {
            f = contact_flux(l, ssl, fl, ss);
        }
        else {
            f = contact_flux(r, ssr, fr, ss);
        }
    }
    to_global(flux, f, basis);
}

static void hlle(double *flux, const double *lp, const double *rp, const double *property,
                 OMat basis) /* src/euler/convective.c:348-389 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double gamma = property[0];
    double sl = sqrt(l.density);
    double sr = sqrt(r.density);
    double w = 1 / (sl + sr);
    double ux = w * ((sl * l.velocity.x) + (sr * r.velocity.x));
    double c2l = gamma * l.pressure / l.density;
    double c2r = gamma * r.pressure / r.density;
    double cl = sqrt(c2l);
    double cr = sqrt(c2r);
    double c2 = w * ((sl * c2l) + (sr * c2r));
    double vj = SQ(w) * sl * sr * SQ(r.velocity.x - l.velocity.x) / 2;
    double c = sqrt(c2 + vj);
    double ssl = fmin(l.velocity.x - cl, ux - c);
    double ssr = fmax(r.velocity.x + cr, ux + c);
    Cons f;
    if (ssl >= 0) {
        f = physical_flux(l);
    }
    else if (ssr <= 0) {
        f = physical_flux(r);
    }
    else {
        f = average_flux(l, r, ssl, ssr);
    }
    to_global(flux, f, basis);
}

/* split Mach/pressure functions shared by ausmd and ausmdv (src/euler/convective.c:399-438,
 * 471-510); `vflag` selects ausmdv's form of the supersonic pressure term */
static void ausm_split(Local l, Local r, double gamma, int vflag, double *c_, double *vpl,
                       double *ppl, double *vmr, double *pmr)
{
    double cl = sqrt(gamma * l.pressure / l.density);
    double cr = sqrt(gamma * r.pressure / r.density);
    double c = fmax(cl, cr);
    double pdl = l.pressure / l.density;
    double pdr = r.pressure / r.density;
    double pds = pdl + pdr;
    double ml = l.velocity.x / c;
    double mr = r.velocity.x / c;
    if (fabs(l.velocity.x) <= c) {
        double al = 2 * pdl / pds;
        double fac = l.velocity.x + c;
        *vpl = (al * SQ(fac) / (4 * c)) + ((1 - al) * (l.velocity.x + fabs(l.velocity.x)) / 2);
        *ppl = (l.pressure * SQ(fac) / (4 * SQ(c))) * (2 - ml);
    }
    else {
        *vpl = (l.velocity.x + fabs(l.velocity.x)) / 2;
        if (vflag) {
            *ppl = l.pressure * (l.velocity.x + fabs(l.velocity.x)) / (2 * l.velocity.x);
        }
        else {
            *ppl = (l.pressure * (l.velocity.x + fabs(l.velocity.x)) / 2) / l.velocity.x;
        }
    }
    if (fabs(r.velocity.x) <= c) {
        double ar = 2 * pdr / pds;
        double fac = r.velocity.x - c;
        *vmr = (-ar * SQ(fac) / (4 * c)) + ((1 - ar) * (r.velocity.x - fabs(r.velocity.x)) / 2);
        *pmr = (r.pressure * SQ(fac) / (4 * SQ(c))) * (2 + mr);
    }
    else {
        *vmr = (r.velocity.x - fabs(r.velocity.x)) / 2;
        if (vflag) {
            *pmr = r.pressure * (r.velocity.x - fabs(r.velocity.x)) / (2 * r.velocity.x);
        }
        else {
            *pmr = (r.pressure * (r.velocity.x - fabs(r.velocity.x)) / 2) / r.velocity.x;
        }
    }
    *c_ = c;
}

static void ausmd(double *flux, const double *lp, const double *rp, const double *property,
                  OMat basis) /* src/euler/convective.c:391-461 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double c;
    double vpl;
    double ppl;
    double vmr;
    double pmr;
    ausm_split(l, r, property[0], 0, &c, &vpl, &ppl, &vmr, &pmr);
    double mf = (vpl * l.density) + (vmr * r.density);
    double amf = fabs(mf);
    double pf = ppl + pmr;
    double hl = (l.energy + l.pressure) / l.density;
    double hr = (r.energy + r.pressure) / r.density;
    Cons f;
    f.density = mf;
    f.momentum.x = (((mf * (l.velocity.x + r.velocity.x)) - (amf * (r.velocity.x - l.velocity.x))) /
                    2) +
                   pf;
    f.momentum.y = ((mf * (l.velocity.y + r.velocity.y)) - (amf * (r.velocity.y - l.velocity.y))) / 2;
    f.momentum.z = ((mf * (l.velocity.z + r.velocity.z)) - (amf * (r.velocity.z - l.velocity.z))) / 2;
    f.energy = ((mf * (hl + hr)) - (amf * (hr - hl))) / 2;
    to_global(flux, f, basis);
}

static void ausmdv(double *flux, const double *lp, const double *rp, const double *property,
                   OMat basis) /* src/euler/convective.c:463-547 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double c;
    double vpl;
    double ppl;
    double vmr;
    double pmr;
    ausm_split(l, r, property[0], 1, &c, &vpl, &ppl, &vmr, &pmr);
    double mf = (vpl * l.density) + (vmr * r.density);
    double amf = fabs(mf);
    double pf = ppl + pmr;
    double nd = ((mf * (l.velocity.x + r.velocity.x)) - (amf * (r.velocity.x - l.velocity.x))) / 2;
    double nv = (vpl * l.momentum.x) + (vmr * r.momentum.x);
    double pmin = fmin(l.pressure, r.pressure);
    double sw = fmin(1.0 / 2, fmin(1, (10.0 * fabs(r.pressure - l.pressure)) / pmin));
    double nm = (((1 + sw) * nv) + ((1 - sw) * nd)) / 2;
    double hl = (l.energy + l.pressure) / l.density;
    double hr = (r.energy + r.pressure) / r.density;
    Cons f;
    f.density = mf;
    f.momentum.x = nm + pf;
    f.momentum.y = ((mf * (l.velocity.y + r.velocity.y)) - (amf * (r.velocity.y - l.velocity.y))) / 2;
    f.momentum.z = ((mf * (l.velocity.z + r.velocity.z)) - (amf * (r.velocity.z - l.velocity.z))) / 2;
    f.energy = ((mf * (hl + hr)) - (amf * (hr - hl))) / 2;
    to_global(flux, f, basis);
}

static void lxf(double *flux, const double *lp, const double *rp, const double *property,
                OMat basis) /* src/euler/convective.c:549-575 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    double gamma = property[0];
    double cl = sqrt(gamma * l.pressure / l.density);
    double cr = sqrt(gamma * r.pressure / r.density);
    double ss = fmax(fabs(l.velocity.x) + cl, fabs(r.velocity.x) + cr);
    Cons j = jump_of(l, r);
    j.density *= ss;
    j.momentum.x *= ss;
    j.momentum.y *= ss;
    j.momentum.z *= ss;
    j.energy *= ss;
    Cons fl = physical_flux(l);
    Cons fr = physical_flux(r);
    Cons f;
    f.density = (fl.density + fr.density - j.density) / 2;
    f.momentum = ov_div(ov_sub(ov_add(fl.momentum, fr.momentum), j.momentum), 2);
    f.energy = (fl.energy + fr.energy - j.energy) / 2;
    to_global(flux, f, basis);
}

static void central(double *flux, const double *lp, const double *rp, const double *property,
                    OMat basis) /* src/euler/convective.c:577-591 */
{
    Local l = to_local(lp, property, basis);
    Local r = to_local(rp, property, basis);
    Cons fl = physical_flux(l);
    Cons fr = physical_flux(r);
    Cons f;
    f.density = (fl.density + fr.density) / 2;
    f.momentum = ov_div(ov_add(fl.momentum, fr.momentum), 2);
    f.energy = (fl.energy + fr.energy) / 2;
    to_global(flux, f, basis);
}

void o_convective(int kind, double *flux, const double *left, const double *right,
                  const double *property, OMat basis)
{
    switch (kind) {
        case O_FLUX_GODUNOV: godunov(flux, left, right, property, basis); return;
        case O_FLUX_ROE: roe(flux, left, right, property, basis); return;
        case O_FLUX_HLL: hll(flux, left, right, property, basis); return;
        case O_FLUX_HLLC: hllc(flux, left, right, property, basis); return;
        case O_FLUX_HLLE: hlle(flux, left, right, property, basis); return;
        case O_FLUX_AUSMD: ausmd(flux, left, right, property, basis); return;
        case O_FLUX_AUSMDV: ausmdv(flux, left, right, property, basis); return;
        case O_FLUX_LXF: lxf(flux, left, right, property, basis); return;
        case O_FLUX_CENTRAL: central(flux, left, right, property, basis); return;
        default: o_fatal("invalid convective flux (%d)", kind);
    }
}

/* name table of src/euler/convective.c:593-623 */
int o_flux_by_name(const char *name)
{
    static const char *names[] = {"godunov", "roe",   "hll", "hllc",   "hlle",
                                  "ausmd",   "ausmdv", "lxf", "central"};
    for (int i = 0; i < 9; i++) {
        if (!strcmp(name, names[i])) {
            return i + 1;
        }
    }
    o_fatal("invalid convective flux (%s)", name);
}

/* src/navierstokes/viscous.c:11-39; gradient = {grad rho, grad u, grad v, grad w, grad p} */
void o_viscous(double *flux, const double *p, const OVec *g, const double *property, OMat basis)
{
    double gamma = property[0];
    double mu = property[1];
    double pr = property[2];
    OMat tau;
    double div = g[1].x + g[2].y + g[3].z;
    tau.x.x = 2 * mu * (g[1].x - (div / 3));
    tau.y.y = 2 * mu * (g[2].y - (div / 3));
    tau.z.z = 2 * mu * (g[3].z - (div / 3));
    tau.y.x = tau.x.y = mu * (g[1].y + g[2].x);
    tau.z.x = tau.x.z = mu * (g[1].z + g[3].x);
    tau.z.y = tau.y.z = mu * (g[2].z + g[3].y);
    double factor = -mu * gamma / ((gamma - 1) * pr * SQ(p[0]));
    OVec heat = ov_mul(factor, ov_sub(ov_mul(p[0], g[4]), ov_mul(p[4], g[0])));
    OVec theta = ov_sub(om_vec(tau, (OVec){p[1], p[2], p[3]}), heat);
    OVec m = om_vec(tau, basis.x);
    flux[0] = 0;
    flux[1] = m.x;
    flux[2] = m.y;
    flux[3] = m.z;
    flux[4] = ov_dot(theta, basis.x);
}

/* ---- boundary conditions: src/euler/boundary.c:7-271, src/navierstokes/boundary.c:6-18 ---- */

static void bc_to_global(double *g, Local l, const double *property, OMat basis)
{ /* src/euler/boundary.c:123-132 */
    l.velocity = ov_div(l.momentum, l.density);
    l.pressure = (property[0] - 1) * (l.energy - (ov_dot(l.momentum, l.velocity) / 2));
    g[0] = fmax(1e-8, l.density);
    OVec v = om_vec(om_transpose(basis), l.velocity);
    g[1] = v.x;
    g[2] = v.y;
    g[3] = v.z;
    g[4] = fmax(1e-8, l.pressure);
}

static void farfield(double *outer, const double *inner_, const double *property, OMat basis,
                     const double *ctx) /* src/euler/boundary.c:134-271 */
{
    Local outerl = to_local(ctx, property, basis);
    Local inner = to_local(inner_, property, basis);
    double gamma = property[0];
    double gm1 = gamma - 1;
    double v2 = ov_norm2(inner.velocity);
    double c2 = gamma * inner.pressure / inner.density;
    double c = sqrt(c2);
    double h = (v2 / 2) + (c2 / gm1);
    OVec u = inner.velocity;
    double evi[5][5] = {
        {h + (c / gm1 * (u.x - c)), -(u.x + (c / gm1)), -u.y, -u.z, 1},
        {(-2 * h) + (4 / gm1 * c2), 2 * u.x, 2 * u.y, 2 * u.z, -2},
        {-2 * u.y * c2 / gm1, 0, 2 * c2 / gm1, 0, 0},
        {-2 * u.z * c2 / gm1, 0, 0, 2 * c2 / gm1, 0},
        {h - (c / gm1 * (u.x + c)), -u.x + (c / gm1), -u.y, -u.z, 1},
    };
    double factor = gm1 / (2 * c2);
    double ui[5] = {inner.density, inner.momentum.x, inner.momentum.y, inner.momentum.z,
                    inner.energy};
    double ug[5] = {outerl.density, outerl.momentum.x, outerl.momentum.y, outerl.momentum.z,
                    outerl.energy};
    double chi[5];
    double chg[5];
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
    double cg = sqrt(gamma * outerl.pressure / outerl.density);
    if (outerl.velocity.x - cg > 0) {
        chg[0] = chi[0];
    }
    if (outerl.velocity.x > 0) {
        chg[1] = chi[1];
        chg[2] = chi[2];
        chg[3] = chi[3];
    }
    if (outerl.velocity.x + cg > 0) {
        chg[4] = chi[4];
    }
    double ev[5][5] = {
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
    outerl.density = ug[0];
    outerl.momentum = (OVec){ug[1], ug[2], ug[3]};
    outerl.energy = ug[4];
    bc_to_global(outer, outerl, property, basis);
}

void o_bc(int kind, double *outer, const double *inner, const double *property, OMat basis,
          const double *ctx)
{
    double gamma = property[0];
    OVec vi = {inner[1], inner[2], inner[3]};
    switch (kind) {
        case O_BC_SYMMETRY: { /* boundary.c:7-20 */
            double vel = ov_dot(vi, basis.x);
            OVec v = ov_sub(vi, ov_mul(2 * vel, basis.x));
            outer[0] = inner[0];
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            outer[4] = inner[4];
            return;
        }
        case O_BC_SUPERSONIC_INFLOW: /* boundary.c:22-32 */
            memcpy(outer, ctx, 5 * sizeof(double));
            return;
        case O_BC_SUPERSONIC_OUTFLOW: /* boundary.c:34-44 */
            memcpy(outer, inner, 5 * sizeof(double));
            return;
        case O_BC_SUBSONIC_INFLOW: { /* boundary.c:46-65 */
            double density = inner[0];
            double c2 = gamma * inner[4] / inner[0];
            double c = sqrt(c2);
            OVec vc = {ctx[1], ctx[2], ctx[3]};
            double vel = ov_dot(ov_sub(vc, vi), basis.x);
            double f1 = density * c;
            outer[4] = (ctx[4] + inner[4] - (f1 * vel)) / 2;
            outer[0] = ctx[0] + ((outer[4] - ctx[4]) / c2);
            double f2 = (ctx[4] - outer[4]) / f1;
            OVec v = ov_sub(vc, ov_mul(f2, basis.x));
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            return;
        }
        case O_BC_SUBSONIC_OUTFLOW: { /* boundary.c:67-84 */
            double density = inner[0];
            double c2 = gamma * inner[4] / inner[0];
            double c = sqrt(c2);
            double pressure = ctx[4];
            outer[4] = pressure;
            outer[0] = inner[0] + ((pressure - inner[4]) / c2);
            double f = (inner[4] - pressure) / (density * c);
            OVec v = ov_sub(vi, ov_mul(f, basis.x));
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            return;
        }
        case O_BC_PRESSURE_OUTFLOW: { /* boundary.c:86-101 */
            double c = sqrt(gamma * inner[4] / inner[0]);
            double vel = ov_dot(vi, basis.x);
            double pressure = (vel / c < 1) ? ctx[4] : inner[4];
            outer[0] = inner[0] * pressure / inner[4];
            outer[1] = inner[1];
            outer[2] = inner[2];
            outer[3] = inner[3];
            outer[4] = pressure;
            return;
        }
        case O_BC_FARFIELD: farfield(outer, inner, property, basis, ctx); return;
        case O_BC_ADIABATIC_WALL: { /* src/navierstokes/boundary.c:6-18 */
            OVec v = ov_mul(-1, vi);
            outer[0] = inner[0];
            outer[1] = v.x;
            outer[2] = v.y;
            outer[3] = v.z;
            outer[4] = inner[4];
            return;
        }
        default: o_fatal("invalid boundary condition (%d)", kind);
    }
}

/* name tables of src/euler/boundary.c:273-297 and src/navierstokes/boundary.c:20-26 */
int o_bc_by_name(const char *name)
{
    if (!strcmp(name, "symmetry") || !strcmp(name, "slipwall")) {
        return O_BC_SYMMETRY;
    }
    if (!strcmp(name, "supersonic inflow")) {
        return O_BC_SUPERSONIC_INFLOW;
    }
    if (!strcmp(name, "supersonic outflow")) {
        return O_BC_SUPERSONIC_OUTFLOW;
    }
    if (!strcmp(name, "subsonic inflow")) {
        return O_BC_SUBSONIC_INFLOW;
    }
    if (!strcmp(name, "subsonic outflow")) {
        return O_BC_SUBSONIC_OUTFLOW;
    }
    if (!strcmp(name, "pressure outflow")) {
        return O_BC_PRESSURE_OUTFLOW;
    }
    if (!strcmp(name, "farfield")) {
        return O_BC_FARFIELD;
    }
    if (!strcmp(name, "adiabatic wall") || !strcmp(name, "wall")) {
        return O_BC_ADIABATIC_WALL;
    }
    o_fatal("invalid boundary condition (%s)", name);
}

/* ---- conversions and time steps: src/euler/create.c:6-44, src/navierstokes/create.c:7-26 ---- */

void o_cons2prim(double *p, const double *u, const double *property)
{
    double gamma = property[0];
    p[0] = fmax(1e-8, u[0]);
    OVec v = ov_div((OVec){u[1], u[2], u[3]}, p[0]);
    p[1] = v.x;
    p[2] = v.y;
    p[3] = v.z;
    p[4] = fmax(1e-8, (gamma - 1) * (u[4] - (ov_dot((OVec){u[1], u[2], u[3]}, v) / 2)));
}

void o_prim2cons(double *u, const double *p, const double *property)
{
    double gamma = property[0];
    u[0] = p[0];
    OVec m = ov_mul(u[0], (OVec){p[1], p[2], p[3]});
    u[1] = m.x;
    u[2] = m.y;
    u[3] = m.z;
    u[4] = (p[4] / (gamma - 1)) + (ov_dot(m, (OVec){p[1], p[2], p[3]}) / 2);
}

double o_euler_timestep(const double *p, const double *property, double volume, OVec proj)
{
    double gamma = property[0];
    double c = sqrt(gamma * p[4] / p[0]);
    double lx = (fabs(p[1]) + c) * proj.x;
    double ly = (fabs(p[2]) + c) * proj.y;
    double lz = (fabs(p[3]) + c) * proj.z;
    return volume / (lx + ly + lz);
}

double o_ns_timestep(const double *p, const double *property, double volume, OVec proj)
{
    double gamma = property[0];
    double mu = property[1];
    double pr = property[2];
    double c = sqrt(gamma * p[4] / p[0]);
    double lx = (fabs(p[1]) + c) * proj.x;
    double ly = (fabs(p[2]) + c) * proj.y;
    double lz = (fabs(p[3]) + c) * proj.z;
    double lv = fmax(4 * mu / (3 * p[0]), gamma * mu / (p[0] * pr));
    return volume / (lx + ly + lz + (4 * lv * ov_norm2(proj) / volume));
}

/* ---- exact Riemann solver: src/euler/riemann.c:1-197 ---- */

static double guess_pressure(const double *l, const double *r, double lc, double rc,
                             const double *gam)
{
    double cup = (l[0] + r[0]) * (lc + rc) / 4;
    double diff = l[1] - r[1];
    double pvrs = fmax((l[4] + r[4]) + (diff * cup), 1e-6) / 2;
    double pmin = fmin(l[4], r[4]);
    double pmax = fmax(l[4], r[4]);
    double qmax = pmax / pmin;
    if (qmax <= 2 && pmin <= pvrs && pvrs <= pmax) {
        return pvrs;
    }
    if (pvrs < pmin) {
        double ratio = pow(l[4] / r[4], gam[1]);
        double guess = ((ratio * l[1] / lc) + (r[1] / rc) + (gam[4] * (ratio - 1))) /
                       ((ratio / lc) + (1 / rc));
        double tl = 1 + (gam[7] * (l[1] - guess) / lc);
        double tr = 1 + (gam[7] * (guess - r[1]) / rc);
        return ((l[4] * pow(tl, gam[3])) + (r[4] * pow(tr, gam[3]))) / 2;
    }
    double sl = sqrt((gam[5] / l[0]) / ((gam[6] * l[4]) + pvrs));
    double sr = sqrt((gam[5] / r[0]) / ((gam[6] * r[4]) + pvrs));
    return ((sl * l[4]) + (sr * r[4]) + diff) / (sl + sr);
}

static void prefun(double *f, double *df, double pold, const double *s, double c,
                   const double *gam)
{
    if (pold <= s[4]) {
        double ratio = pold / s[4];
        *f = gam[4] * c * (pow(ratio, gam[1]) - 1);
        *df = (1 / (s[0] * c)) * pow(ratio, -gam[2]);
    }
    else {
        double a = gam[5] / s[0];
        double b = gam[6] * s[4];
        double root = sqrt(a / (b + pold));
        *f = (pold - s[4]) * root;
        *df = (1 - ((pold - s[4]) / (b + pold) / 2)) * root;
    }
}

static void sample(const double *l, const double *r, double lc, double rc, double su, double sp,
                   const double *gam, double loc, double *o)
{
    o[2] = o[3] = 0;
    if (loc <= su) {
        if (sp <= l[4]) {
            if (loc <= l[1] - lc) {
                memcpy(o, l, 5 * sizeof(double));
                return;
            }
            double sc = lc * pow(sp / l[4], gam[1]);
            if (loc > su - sc) {
                o[0] = l[0] * pow(sp / l[4], 1 / gam[0]);
                o[1] = su;
                o[4] = sp;
                return;
            }
            double fc = gam[5] * (lc + (gam[7] * (l[1] - loc)));
            o[0] = l[0] * pow(fc / lc, gam[4]);
            o[1] = gam[5] * (lc + (gam[7] * l[1]) + loc);
            o[4] = l[4] * pow(fc / lc, gam[3]);
            return;
        }
        double ratio = sp / l[4];
        if (loc <= l[1] - (lc * sqrt((gam[2] * ratio) + gam[1]))) {
            memcpy(o, l, 5 * sizeof(double));
            return;
        }
        o[0] = l[0] * (ratio + gam[6]) / ((ratio * gam[6]) + 1);
        o[1] = su;
        o[4] = sp;
        return;
    }
    if (sp > r[4]) {
        double ratio = sp / r[4];
        if (loc >= r[1] + (rc * sqrt((gam[2] * ratio) + gam[1]))) {
            memcpy(o, r, 5 * sizeof(double));
            return;
        }
        o[0] = r[0] * (ratio + gam[6]) / ((ratio * gam[6]) + 1);
        o[1] = su;
        o[4] = sp;
        return;
    }
    if (loc >= r[1] + rc) {
        memcpy(o, r, 5 * sizeof(double));
        return;
    }
    double sc = rc * pow(sp / r[4], gam[1]);
    if (loc <= su + sc) {
        o[0] = r[0] * pow(sp / r[4], 1 / gam[0]);
        o[1] = su;
        o[4] = sp;
        return;
    }
    double fc = gam[5] * (rc - (gam[7] * (r[1] - loc)));
    o[0] = r[0] * pow(fc / rc, gam[4]);
    o[1] = gam[5] * (-rc + (gam[7] * r[1]) + loc);
    o[4] = r[4] * pow(fc / rc, gam[3]);
}

void o_riemann(const double *l, const double *r, double gamma, double loc, double *out)
{
    double gam[9] = {gamma,
                     (gamma - 1) / (2 * gamma),
                     (gamma + 1) / (2 * gamma),
                     2 * gamma / (gamma - 1),
                     2 / (gamma - 1),
                     2 / (gamma + 1),
                     (gamma - 1) / (gamma + 1),
                     (gamma - 1) / 2,
                     gamma - 1};
    double lc = sqrt(gamma * l[4] / l[0]);
    double rc = sqrt(gamma * r[4] / r[0]);
    if (gam[4] * (lc + rc) <= (r[1] - l[1])) {
        o_fatal("vacuum state detected in riemann solver");
    }
    double diff = r[1] - l[1];
    double pold = guess_pressure(l, r, lc, rc, gam);
    double su = 0;
    double sp = 0;
    int done = 0;
    for (int iter = 0; iter < 20; iter++) {
        double lf;
        double ldf;
        double rf;
        double rdf;
        prefun(&lf, &ldf, pold, l, lc, gam);
        prefun(&rf, &rdf, pold, r, rc, gam);
        double pnew = pold - ((lf + rf + diff) / (ldf + rdf));
        if (2 * fabs((pnew - pold) / (pnew + pold)) <= 1e-6) {
            su = (l[1] + r[1] + rf - lf) / 2;
            sp = pnew;
            done = 1;
            break;
        }
        pold = fmax(1e-6, pnew);
    }
    if (!done) {
        o_fatal("newton solver did not converge (20)");
    }
    sample(l, r, lc, rc, su, sp, gam, loc, out);
    if (loc <= su) {
        out[2] = l[2];
        out[3] = l[3];
    }
    else {
        out[2] = r[2];
        out[3] = r[3];
    }
}
