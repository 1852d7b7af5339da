// mom6h_model.hip -- one run of the split-RK2 driver on one GPU.
//
// Replays app/rk2_serial_driver.F90:218-399 with every module on the device:
//   predictor  transports -> hor_visc -> CorAdCalc -> velocity update ->
//              vert_visc -> [btstep] -> continuity
//   corrector  the same on the predicted state, then the final continuity
// The barotropic chain (btstep predictor -> corrector) reads only eta/ubt/vbt
// and writes only ubt_av/vbt_av/eta_av/eta, none of which the 3D stages touch
// (rk2_serial_driver.F90:285,365), so it runs on a second stream and joins the
// main stream at the end of the step.
//
// Two copies of the reference driver are avoided without changing any value:
//   * h = h0 at the top of the step (:222-228): the predictor stages read h0
//     directly and the predictor continuity writes the whole h_pred field
//     (compute domain updated, halo = h0);
//   * htmp = h before the final continuity (:373-379): the predictor
//     continuity writes straight into htmp and the final one writes h from it.
// For periodic grids the 3-wide halos are refilled at the exchange points of
// app/rk2_mpi_serial_driver.F90 (uh/vh after transports, up/vp before the
// predictor continuity, h after it, u/v/h/eta/ubt/vbt at the end of the step).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mom6h_model.h"

using namespace mom6h;

namespace {

const double kG_EARTH = 9.80;     // src/common/mom6_types.F90:80
const double kDepth = 4000.0;     // mom6_barotropic_serial.F90:76

}  // namespace

namespace mom6h {
const char *kStageNames[ST_COUNT] = {"transports", "hor_visc",   "coriolis", "vel_update", "vert_visc",
                                     "continuity", "halo",       "btstep",   "momentum"};

}  // namespace mom6h

namespace {

int add_field(mom6h_model *m, const char *name, double **p, size_t n, bool zero)
{
    int rc = dev_alloc(p, n, zero);
    if (rc) return rc;
    m->fields.push_back(Field{name, (void **)p, n, 8, true});
    return MOM6H_OK;
}

int add_field(mom6h_model *m, const char *name, float **p, size_t n, bool zero)
{
    int rc = dev_alloc_bytes((void **)p, n * sizeof(float), zero);
    if (rc) return rc;
    m->fields.push_back(Field{name, (void **)p, n, 4, true});
    return MOM6H_OK;
}

const Field *find_field(const mom6h_model *m, const char *name)
{
    for (auto &f : m->fields)
        if (f.name == name) return &f;
    return nullptr;
}

// init_ocean_grid (src/common/mom6_types.F90:130-258): uniform spacing dx.
void host_grid(mom6h_model *m)
{
    const size_t n2 = (size_t)m->d.s2;
    const double dx_m = m->c.dx_km * 1000.0;
    auto fill = [&](std::vector<double> &a, double v) { a.assign(n2, v); };
    fill(m->areaT, dx_m * dx_m);
    fill(m->IareaT, 1.0 / (dx_m * dx_m));
    fill(m->dxT, dx_m); fill(m->dyT, dx_m); fill(m->IdxT, 1.0 / dx_m); fill(m->IdyT, 1.0 / dx_m);
    fill(m->dxCu, dx_m); fill(m->dyCu, dx_m); fill(m->dy_Cu, dx_m);
    fill(m->IdxCu, 1.0 / dx_m); fill(m->IdyCu, 1.0 / dx_m);
    fill(m->dxCv, dx_m); fill(m->dyCv, dx_m); fill(m->IdxCv, 1.0 / dx_m); fill(m->IdyCv, 1.0 / dx_m);
    fill(m->IareaBu, 1.0 / (dx_m * dx_m)); fill(m->IareaCu, 1.0 / (dx_m * dx_m));
    fill(m->IareaCv, 1.0 / (dx_m * dx_m));
    fill(m->dxBu, dx_m); fill(m->dyBu, dx_m); fill(m->IdxBu, 1.0 / dx_m); fill(m->IdyBu, 1.0 / dx_m);
}

std::vector<double> download(const double *d_p, size_t n)
{
    std::vector<double> h(n);
    if (hipMemcpy(h.data(), d_p, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) h.clear();
    return h;
}

template <typename T>
int fill_halos(mom6h_model *m, std::initializer_list<std::pair<T *, int>> arrays)
{
    if (!m->halo.px && !m->halo.py) return MOM6H_OK;
    for (auto &a : arrays) {
        int rc = halo_fill(m->d, m->halo, a.first, a.second, m->s0);
        if (rc) return rc;
    }
    return MOM6H_OK;
}

int fill_halos(mom6h_model *m, std::initializer_list<std::pair<double *, int>> arrays)
{
    return fill_halos<double>(m, arrays);
}

int fill_halos(mom6h_model *m, std::initializer_list<std::pair<float *, int>> arrays)
{
    return fill_halos<float>(m, arrays);
}

#define RC(expr)               \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

struct StageTimer {
    mom6h_model *m;
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(mom6h_model *m_, int st, hipStream_t s_) : m(m_), stage(st), s(s_)
    {
        if (m->profile) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, s);
        }
    }
    ~StageTimer()
    {
        if (m->profile) {
            (void)hipEventRecord(b, s);
            m->ev.push_back(a);
            m->ev.push_back(b);
            m->ev.push_back((hipEvent_t)(intptr_t)stage);
        }
    }
};

// The coupled form's targets: uhbt = Datu*ubt_av, vhbt = Datv*vbt_av over the
// data domain, from the btstep call just made (ubt_av, vbt_av vanish outside
// btstep's ranges, so the boundary faces of a closed domain get no transport)
__global__ void k_coupled_targets(long long n, const double *__restrict__ Datu, const double *__restrict__ ubt_av,
                                  const double *__restrict__ Datv, const double *__restrict__ vbt_av,
                                  double *__restrict__ uhbt, double *__restrict__ vhbt)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    uhbt[q] = Datu[q] * ubt_av[q];
    vhbt[q] = Datv[q] * vbt_av[q];
}

// One btstep call of the coupled form on the 3D stream, then the targets of
// the continuity call that follows (predictor: eta, ubt, vbt -> averages;
// corrector: the averages -> eta, as in the reference driver, :285 and :365)
int coupled_btstep(mom6h_model *m, bool corrector)
{
    hipStream_t s0 = m->s0;
    {
        StageTimer t(m, ST_BTSTEP, s0);
        if (!corrector) RC(mom6h_btstep(m->bt_cs, m->eta, m->ubt, m->vbt, m->ubt_av, m->vbt_av, m->eta_av, s0));
        else RC(mom6h_btstep(m->bt_cs, m->eta_av, m->ubt_av, m->vbt_av, m->ubt_av, m->vbt_av, m->eta, s0));
    }
    const long long n = m->d.s2;
    k_coupled_targets<<<(unsigned)((n + 255) / 256), 256, 0, s0>>>(n, m->bt_cs->Datu, m->ubt_av, m->bt_cs->Datv,
                                                                    m->vbt_av, m->uhbt, m->vhbt);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

// The barotropic chain of one step on stream s1 (:285, :365): predictor and
// corrector btstep, then (periodic grids) the end-of-step refill of the 2D
// fields it owns.  It never waits for the 3D stream.
int enqueue_barotropic(mom6h_model *m)
{
    hipStream_t s1 = m->serialize ? m->s0 : m->s1;
    {
        StageTimer t(m, ST_BTSTEP, s1);
        RC(mom6h_btstep(m->bt_cs, m->eta, m->ubt, m->vbt, m->ubt_av, m->vbt_av, m->eta_av, s1));
        RC(mom6h_btstep(m->bt_cs, m->eta_av, m->ubt_av, m->vbt_av, m->ubt_av, m->vbt_av, m->eta, s1));
    }
    if (m->halo.px || m->halo.py)
        for (double *a : {m->eta, m->ubt, m->vbt}) RC(halo_fill(m->d, m->halo, a, 1, s1));
    return MOM6H_OK;
}


// transports + hor_visc + CorAdCalc + velocity update of one phase
int momentum_phase(mom6h_model *m, const double *u, const double *v, const double *h, double *ou,
                   double *ov, double fdt)
{
    const Dims &d = m->d;
    const int nk = m->c.nk;
    hipStream_t s0 = m->s0;
    if (fused_ok(m)) {
        StageTimer t(m, ST_MOMENTUM, s0);
        MomArgs a;
        a.d = d; a.g = m->ug; a.Kh = m->c.Kh; a.h_neglect = 1.0e-3; a.fdt = fdt; a.no_slip = m->c.no_slip;
        a.u = u; a.v = v; a.h = h; a.bu = m->u; a.bv = m->v; a.ou = ou; a.ov = ov;
        a.mask2dT = m->mask2dT; a.mask2dBu = m->mask2dBu; a.CoriolisBu = m->CoriolisBu;
        a.Ah = m->c.Ah; a.del2u = m->del2u; a.del2v = m->del2v; a.nkz = m->momentum_splits;
        return launch_momentum(a, s0, m->momentum_tile);
    }
    { StageTimer t(m, ST_TRANSPORTS, s0);
      RC(mom6h_transports(d.isd, d.ied, d.jsd, d.jed, nk, u, v, h, m->cor->dyCu, m->cor->dxCv, m->uh, m->vh, s0)); }
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{m->uh, nk}, {m->vh, nk}})); }
    { StageTimer t(m, ST_HOR_VISC, s0); RC(mom6h_hor_visc(m->hv, u, v, h, m->diffu, m->diffv, s0)); }
    { StageTimer t(m, ST_CORIOLIS, s0); RC(mom6h_coradcalc(m->cor, u, v, h, m->uh, m->vh, m->CAu, m->CAv, s0)); }
    { StageTimer t(m, ST_VEL_UPDATE, s0);
      RC(mom6h_vel_update(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, nk, fdt, m->u, m->v, m->CAu,
                          m->diffu, m->CAv, m->diffv, ou, ov, s0)); }
    return MOM6H_OK;
}

// continuity_PPM of one phase: u, hin -> hout, and with the meridional option
// the meridional pass after the zonal one, on the zonal result with its
// periodic halos refilled.
int continuity_phase(mom6h_model *m, const double *u, const double *v, const double *hin, double *hout,
                     double dt)
{
    hipStream_t s0 = m->s0;
    StageTimer t(m, ST_CONTINUITY, s0);
    const double *uhbt = m->c.coupled ? m->uhbt : nullptr, *vhbt = m->c.coupled ? m->vhbt : nullptr;
    if (!m->c.meridional)
        return mom6h_continuity_ppm(m->cont, u, hin, hout, m->uh, dt, uhbt, &m->bt, m->u_cor, m->du_cor, 1, s0);
    RC(mom6h_continuity_ppm(m->cont, u, hin, m->hx, m->uh, dt, uhbt, &m->bt, m->u_cor, m->du_cor, 1, s0));
    RC(fill_halos(m, {{m->hx, m->c.nk}}));
    return mom6h_continuity_ppm_y(m->cont, v, m->hx, hout, m->vh, dt, vhbt, &m->bt_v, m->v_cor, m->dv_cor, 1,
                                  s0);
}

// ---- the mixed-precision form: FP32 layer fields and stencils, FP64
// vertical solve, barotropic chain and 2D fields
int momentum_phase_f32(mom6h_model *m, const float *u, const float *v, const float *h, float *ou, float *ov,
                       double fdt)
{
    StageTimer t(m, ST_MOMENTUM, m->s0);
    MomArgsT<float> a;
    a.d = m->d; a.g = m->ug; a.Kh = (float)m->c.Kh; a.h_neglect = 1.0e-3f; a.fdt = (float)fdt;
    a.no_slip = m->c.no_slip;
    a.u = u; a.v = v; a.h = h; a.bu = m->f.u; a.bv = m->f.v; a.ou = ou; a.ov = ov;
    a.mask2dT = m->mask2dT; a.mask2dBu = m->mask2dBu; a.CoriolisBu = m->CoriolisBu;
    a.Ah = (float)m->c.Ah; a.del2u = m->f.del2u; a.del2v = m->f.del2v; a.nkz = m->momentum_splits;
    return launch_momentum(a, m->s0, m->momentum_tile);
}

int continuity_phase_f32(mom6h_model *m, const float *u, const float *v, const float *hin, float *hout, double dt)
{
    hipStream_t s0 = m->s0;
    StageTimer t(m, ST_CONTINUITY, s0);
    if (!m->c.meridional)
        return continuity_f32(m->cont, u, hin, hout, m->f.uh, dt, &m->bt, m->f.bt_h_u, m->f.u_cor, m->du_cor, 1, s0);
    RC(continuity_f32(m->cont, u, hin, m->f.hx, m->f.uh, dt, &m->bt, m->f.bt_h_u, m->f.u_cor, m->du_cor, 1, s0));
    RC(fill_halos(m, {{m->f.hx, m->c.nk}}));
    return continuity_y_f32(m->cont, v, m->f.hx, hout, m->f.vh, dt, &m->bt_v, m->f.bt_h_v, m->f.v_cor, m->dv_cor,
                            1, s0);
}

int enqueue_step_f32(mom6h_model *m)
{
    const double dt = m->c.dt;
    const int nk = m->c.nk;
    hipStream_t s0 = m->s0;
    auto &f = m->f;
    // ---- predictor (h = h0 throughout)
    RC(momentum_phase_f32(m, f.u, f.v, f.h0, f.up, f.vp, dt));
    { StageTimer t(m, ST_VERT_VISC, s0);
      RC(vert_visc_f32(m->vv, f.up, f.vp, f.h0, dt, m->taux, m->tauy, m->ray_u, m->ray_v, s0)); }
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{f.up, nk}, {f.vp, nk}})); }
    RC(continuity_phase_f32(m, f.up, f.vp, f.h0, f.htmp, dt));
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{f.htmp, nk}})); }
    // ---- corrector (h = h_pred, held in htmp)
    RC(momentum_phase_f32(m, f.up, f.vp, f.htmp, f.u, f.v, 0.5 * dt));
    { StageTimer t(m, ST_VERT_VISC, s0);
      RC(vert_visc_f32(m->vv, f.u, f.v, f.htmp, 0.5 * dt, m->taux, m->tauy, m->ray_u, m->ray_v, s0)); }
    RC(continuity_phase_f32(m, f.u, f.v, f.htmp, f.h, 0.5 * dt));
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{f.u, nk}, {f.v, nk}, {f.h, nk}})); }
    return MOM6H_OK;
}

// The predictor's diagnostics (rk2_serial_driver.F90:300-308): the vertical
// sums of diffu*h and diffv*h (h = the predicted thickness) posted in turn
// through the one shared work array, each followed by its statistics
// (diag_out[4..7] and [8..11]).  The fused stencil keeps diffu on chip: the
// predictor's viscous tendencies are formed here by the module kernel.
int post_predictor_diagnostics(mom6h_model *m)
{
    const Dims &d = m->d;
    hipStream_t s0 = m->s0;
    if (fused_ok(m)) RC(mom6h_hor_visc(m->hv, m->u, m->v, m->h0, m->diffu, m->diffv, s0));
    RC(mom6h_diag_product_sum(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, d.nk, m->diffu, m->htmp, 0,
                              m->diag_work2d, s0));
    RC(mom6h_diag_stats(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, 1, m->diag_work2d, m->diag_red,
                        m->diag_out + 4, s0));
    RC(mom6h_diag_product_sum(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, d.nk, m->diffv, m->htmp, 1,
                              m->diag_work2d, s0));
    return mom6h_diag_stats(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, 1, m->diag_work2d, m->diag_red,
                            m->diag_out + 8, s0);
}

// vert_visc_coef + vert_visc_apply of one phase: Rayleigh drag in the bottom
// layer (2D) or at every level (3D fields)
int vert_visc_phase(mom6h_model *m, double *u, double *v, const double *h, double dt)
{
    if (m->c.ray3d)
        return mom6h_vert_visc_ray3d(m->vv, u, v, h, dt, m->taux, m->tauy, m->ray_u3d, m->ray_v3d, m->s0);
    return mom6h_vert_visc(m->vv, u, v, h, dt, m->taux, m->tauy, m->ray_u, m->ray_v, m->s0);
}

// The 3D part of one RK2 step (rk2_serial_driver.F90:218-399) on stream s0.
int enqueue_step(mom6h_model *m)
{
    if (m->mixed) return enqueue_step_f32(m);
    const double dt = m->c.dt;
    const int nk = m->c.nk;
    hipStream_t s0 = m->s0;

    // ---- predictor (h = h0 throughout)
    RC(momentum_phase(m, m->u, m->v, m->h0, m->up, m->vp, dt));
    { StageTimer t(m, ST_VERT_VISC, s0); RC(vert_visc_phase(m, m->up, m->vp, m->h0, dt)); }
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{m->up, nk}, {m->vp, nk}})); }
    if (m->c.coupled) RC(coupled_btstep(m, false));
    RC(continuity_phase(m, m->up, m->vp, m->h0, m->htmp, dt));
    if (m->diag) RC(post_predictor_diagnostics(m));
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{m->htmp, nk}})); }

    // ---- corrector (h = h_pred, held in htmp)
    RC(momentum_phase(m, m->up, m->vp, m->htmp, m->u, m->v, 0.5 * dt));
    { StageTimer t(m, ST_VERT_VISC, s0); RC(vert_visc_phase(m, m->u, m->v, m->htmp, 0.5 * dt)); }
    if (m->c.coupled) RC(coupled_btstep(m, true));
    RC(continuity_phase(m, m->u, m->v, m->htmp, m->h, 0.5 * dt));
    { StageTimer t(m, ST_HALO, s0); RC(fill_halos(m, {{m->u, nk}, {m->v, nk}, {m->h, nk}})); }
    if (m->c.coupled && (m->halo.px || m->halo.py))   // the barotropic chain's end-of-step refill
        for (double *a : {m->eta, m->ubt, m->vbt}) RC(halo_fill(m->d, m->halo, a, 1, s0));
    return MOM6H_OK;
}

}  // namespace

extern "C" {

int mom6h_model_create(mom6h_model **out, const mom6h_config *c)
{
    MOM6H_REQUIRE(out && c, "mom6h_model_create: NULL argument");
    MOM6H_REQUIRE(c->ni >= 10 && c->nj >= 10 && c->nk >= 1, "mom6h_model_create: ni, nj >= 10 and nk >= 1 required");
    MOM6H_REQUIRE(c->bt_nsteps >= 1 && c->dt > 0.0 && c->dx_km > 0.0, "mom6h_model_create: bad bt_nsteps/dt/dx");
    MOM6H_REQUIRE(!(c->periodic_y && c->walls_y), "mom6h_model_create: periodic_y and walls_y exclude each other");
    MOM6H_REQUIRE(c->coriolis_scheme >= 1 && c->coriolis_scheme <= 3, "mom6h_model_create: bad coriolis_scheme");
    MOM6H_REQUIRE(c->precision == 0 || c->precision == 1, "mom6h_model_create: precision must be 0 (FP64) or 1 (mixed)");
    MOM6H_REQUIRE(c->precision == 0 || (c->coriolis_scheme == 1 && c->laplacian && c->nk <= 80 &&
                                        !c->smagorinsky && !c->leith && !c->better_bound && !c->coupled &&
                                        !c->ray3d),
                  "mom6h_model_create: the mixed-precision form runs the fused stencils: Sadourny Coriolis, "
                  "Laplacian (+ biharmonic) viscosity, nk <= 80");
    auto *m = new mom6h_model();
    m->c = *c;
    m->mixed = c->precision == 1;
    // mom6_types.F90:145-148
    m->d = make_dims(1, c->ni + 2, 1, c->nj + 2, 4, c->ni - 1, 4, c->nj - 1, c->nk);
    m->halo = Halo{c->periodic_x, c->periodic_y};
    const size_t n2 = (size_t)m->d.s2, n3 = n2 * c->nk;
    int rc = 0;
#define F3(x) rc = rc ? rc : add_field(m, #x, &m->x, n3, true)
#define G3(x) rc = rc ? rc : add_field(m, #x, &m->f.x, n3, true)
#define F2(x) rc = rc ? rc : add_field(m, #x, &m->x, n2, true)
    if (!m->mixed) {
        F3(u); F3(v); F3(h); F3(h0); F3(htmp); F3(uh); F3(vh); F3(CAu); F3(CAv); F3(up); F3(vp);
        F3(diffu); F3(diffv); F3(u_cor);
    } else {
        G3(u); G3(v); G3(h); G3(h0); G3(htmp); G3(uh); G3(vh); G3(up); G3(vp); G3(u_cor);
    }
    F2(eta); F2(ubt); F2(vbt); F2(ubt_av); F2(vbt_av); F2(eta_av); F2(du_cor);
    F2(taux); F2(tauy); F2(ray_u); F2(ray_v); F2(CoriolisBu); F2(mask2dT); F2(mask2dCu);
    F2(mask2dCv); F2(mask2dBu);
    if (c->coupled) { F2(uhbt); F2(vhbt); }
    if (c->ray3d) { F3(ray_u3d); F3(ray_v3d); }
#undef F3
#undef G3
#undef F2
    // BT_cont: six 2D fields (FP64) and the 3D marginal thickness (FP32 in the mixed form)
    auto bt_fields = [&](mom6h_bt_cont &bt, const char *const names[7], float **h3_f32) {
        if (rc) return;
        if (!m->mixed) {
            rc = mom6h_bt_cont_alloc(&bt, 1, c->ni + 2, 1, c->nj + 2, c->nk);
        } else {
            for (double **p : {&bt.FA_u_EE, &bt.FA_u_E0, &bt.FA_u_W0, &bt.FA_u_WW, &bt.uBT_WW, &bt.uBT_EE})
                rc = rc ? rc : dev_alloc(p, n2, true);
        }
        if (rc) return;
        double **p2[6] = {&bt.FA_u_EE, &bt.FA_u_E0, &bt.FA_u_W0, &bt.FA_u_WW, &bt.uBT_WW, &bt.uBT_EE};
        for (int q = 0; q < 6; ++q) m->fields.push_back(Field{names[q], (void **)p2[q], n2, 8, false});
        if (!m->mixed) m->fields.push_back(Field{names[6], (void **)&bt.h_u, n3, 8, false});
        else rc = add_field(m, names[6], h3_f32, n3, true);
    };
    if (!rc && c->biharmonic) {   // zero outside the Del2 ranges, for good
        if (!m->mixed) {
            rc = add_field(m, "del2u", &m->del2u, n3, true);
            rc = rc ? rc : add_field(m, "del2v", &m->del2v, n3, true);
        } else {
            rc = add_field(m, "del2u", &m->f.del2u, n3, true);
            rc = rc ? rc : add_field(m, "del2v", &m->f.del2v, n3, true);
        }
    }
    static const char *const bt_u_names[7] = {"FA_u_EE", "FA_u_E0", "FA_u_W0", "FA_u_WW", "uBT_WW", "uBT_EE", "BT_h_u"};
    static const char *const bt_v_names[7] = {"FA_v_NN", "FA_v_N0", "FA_v_S0", "FA_v_SS", "vBT_SS", "vBT_NN", "BT_h_v"};
    bt_fields(m->bt, bt_u_names, &m->f.bt_h_u);
    if (!rc && c->meridional) {
        if (!m->mixed) {
            rc = add_field(m, "hx", &m->hx, n3, true);
            rc = rc ? rc : add_field(m, "v_cor", &m->v_cor, n3, true);
        } else {
            rc = add_field(m, "hx", &m->f.hx, n3, true);
            rc = rc ? rc : add_field(m, "v_cor", &m->f.v_cor, n3, true);
        }
        rc = rc ? rc : add_field(m, "dv_cor", &m->dv_cor, n2, true);
        bt_fields(m->bt_v, bt_v_names, &m->f.bt_h_v);
    }
    if (!rc && hipStreamCreateWithFlags(&m->s0, hipStreamNonBlocking) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "stream creation failed");
    if (!rc && hipStreamCreateWithFlags(&m->s1, hipStreamNonBlocking) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "stream creation failed");
    if (!rc && hipEventCreateWithFlags(&m->e_start, hipEventDisableTiming) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "event creation failed");
    if (!rc && hipEventCreateWithFlags(&m->e_bt, hipEventDisableTiming) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "event creation failed");
    if (rc) {
        mom6h_model_destroy(m);
        return rc;
    }
    host_grid(m);
    // Rayleigh drag in the bottom layer (rk2_serial_driver.F90:534-539) and
    // the reference's default grid inputs; callers overwrite them by upload.
    {
        std::vector<double> one(n2, 1.0), ray(n2, 1.0e-4);
        (void)hipMemcpy(m->ray_u, ray.data(), n2 * sizeof(double), hipMemcpyHostToDevice);
        (void)hipMemcpy(m->ray_v, ray.data(), n2 * sizeof(double), hipMemcpyHostToDevice);
        for (double *p : {m->mask2dT, m->mask2dCu, m->mask2dCv, m->mask2dBu})
            (void)hipMemcpy(p, one.data(), n2 * sizeof(double), hipMemcpyHostToDevice);
    }
    *out = m;
    return MOM6H_OK;
}

int mom6h_model_field(mom6h_model *m, const char *name, void **d_ptr, long *count)
{
    MOM6H_REQUIRE(m && name, "mom6h_model_field: NULL argument");
    if (const Field *f = find_field(m, name)) {
        if (d_ptr) *d_ptr = *f->p;
        if (count) *count = (long)f->n;
        return MOM6H_OK;
    }
    if (std::string(name) == "vv_taux_bot" || std::string(name) == "vv_tauy_bot" ||
        std::string(name) == "vv_a_u" || std::string(name) == "vv_h_u") {
        MOM6H_REQUIRE(m->vv, "mom6h_model_field: %s exists after prepare", name);
        const size_t n2 = (size_t)m->d.s2;
        double *p = std::string(name) == "vv_taux_bot" ? m->vv->taux_bot
                    : std::string(name) == "vv_tauy_bot" ? m->vv->tauy_bot
                    : std::string(name) == "vv_a_u" ? m->vv->a_u : m->vv->h_u;
        const size_t n = std::string(name) == "vv_a_u" ? n2 * (m->c.nk + 1)
                         : std::string(name) == "vv_h_u" ? n2 * m->c.nk : n2;
        if (d_ptr) *d_ptr = p;
        if (count) *count = (long)n;
        return MOM6H_OK;
    }
    return set_error(MOM6H_EINVAL, "mom6h_model_field: unknown field '%s'", name);
}

int mom6h_model_field_bytes(mom6h_model *m, const char *name, int *elem_bytes)
{
    MOM6H_REQUIRE(m && name && elem_bytes, "mom6h_model_field_bytes: NULL argument");
    const Field *f = find_field(m, name);
    long n = 0;
    if (!f) RC(mom6h_model_field(m, name, nullptr, &n));   // the vert_visc arrays: FP64
    *elem_bytes = f ? f->bytes : 8;
    return MOM6H_OK;
}

int mom6h_model_upload(mom6h_model *m, const char *name, const double *src, long count)
{
    void *p = nullptr;
    long n = 0;
    int bytes = 8;
    RC(mom6h_model_field(m, name, &p, &n));
    RC(mom6h_model_field_bytes(m, name, &bytes));
    MOM6H_REQUIRE(src && count == n, "mom6h_model_upload(%s): expected %ld values, got %ld", name, n, count);
    MOM6H_CHECK(hipDeviceSynchronize());
    if (bytes == 8) {
        MOM6H_CHECK(hipMemcpy(p, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp((size_t)n);
        for (size_t q = 0; q < (size_t)n; ++q) tmp[q] = (float)src[q];
        MOM6H_CHECK(hipMemcpy(p, tmp.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    }
    return MOM6H_OK;
}

int mom6h_model_fill(mom6h_model *m, const char *name, const double *d_src, long count)
{
    void *p = nullptr;
    long n = 0;
    int bytes = 8;
    RC(mom6h_model_field(m, name, &p, &n));
    RC(mom6h_model_field_bytes(m, name, &bytes));
    MOM6H_REQUIRE(d_src && count == n, "mom6h_model_fill(%s): expected %ld values, got %ld", name, n, count);
    MOM6H_CHECK(hipDeviceSynchronize());
    if (bytes == 8) MOM6H_CHECK(hipMemcpy(p, d_src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice));
    else RC(narrow_to_f32(d_src, (float *)p, n, m->s0));
    MOM6H_CHECK(hipStreamSynchronize(m->s0));
    return MOM6H_OK;
}

int mom6h_model_download(mom6h_model *m, const char *name, double *dst, long count)
{
    void *p = nullptr;
    long n = 0;
    int bytes = 8;
    RC(mom6h_model_field(m, name, &p, &n));
    RC(mom6h_model_field_bytes(m, name, &bytes));
    MOM6H_REQUIRE(dst && count == n, "mom6h_model_download(%s): expected %ld values, got %ld", name, n, count);
    MOM6H_CHECK(hipDeviceSynchronize());
    if (bytes == 8) {
        MOM6H_CHECK(hipMemcpy(dst, p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp((size_t)n);
        MOM6H_CHECK(hipMemcpy(tmp.data(), p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < (size_t)n; ++q) dst[q] = (double)tmp[q];
    }
    return MOM6H_OK;
}

int mom6h_model_prepare(mom6h_model *m)
{
    MOM6H_REQUIRE(m != nullptr, "mom6h_model_prepare: NULL model");
    MOM6H_REQUIRE(!m->prepared, "mom6h_model_prepare: already prepared");
    MOM6H_CHECK(hipDeviceSynchronize());
    const Dims &d = m->d;
    const size_t n2 = (size_t)d.s2;
    const int nk = m->c.nk;
    std::vector<double> mT = download(m->mask2dT, n2), mCu = download(m->mask2dCu, n2);
    std::vector<double> mCv = download(m->mask2dCv, n2), mBu = download(m->mask2dBu, n2);
    std::vector<double> fq = download(m->CoriolisBu, n2);
    MOM6H_REQUIRE(!mT.empty() && !mCu.empty() && !mCv.empty() && !mBu.empty() && !fq.empty(),
                  "mom6h_model_prepare: download of grid inputs failed");
    // continuity_init (mom6_continuity_serial.F90:53-78)
    RC(mom6h_continuity_init(&m->cont, d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, nk,
                             m->IareaT.data(), m->IdxT.data(), m->dy_Cu.data(), mT.data(),
                             m->c.monotonic, m->dxT.data(), m->dxCu.data(), mCu.data()));
    if (m->c.meridional)
        RC(mom6h_continuity_set_meridional(m->cont, m->IdyT.data(), m->dxCv.data(), m->dyT.data(), m->dyCv.data(),
                                           mCv.data()));
    // coriolis_init (mom6_coriolis_serial.F90:48-84); the mixed form uses
    // the fused stencil only
    if (!m->mixed)
        RC(mom6h_coriolis_init(&m->cor, d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, nk,
                               m->areaT.data(), m->IareaBu.data(), fq.data(), mBu.data(), m->dyCv.data(),
                               m->dxCu.data(), m->dyCu.data(), m->dxCv.data(), m->IdxCu.data(),
                               m->IdyCv.data(), m->c.coriolis_scheme));
    // barotropic_init (mom6_barotropic_serial.F90:63-137); closed v faces carry no transport
    {
        std::vector<double> Datu(n2), Datv(n2), g(n2, kG_EARTH), rem(n2, 0.999), f4(4 * n2);
        for (size_t q = 0; q < n2; ++q) {
            Datu[q] = kDepth * m->dyCu[q];
            Datv[q] = kDepth * m->dxCv[q];
            if (m->c.walls_y) Datv[q] = Datv[q] * mCv[q];
            for (int cc = 0; cc < 4; ++cc) f4[4 * q + cc] = 0.25 * fq[q];
        }
        RC(mom6h_barotropic_init(&m->bt_cs, d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec,
                                 m->c.bt_nsteps, m->c.dt, 0.2, 0, Datu.data(), Datv.data(), g.data(),
                                 g.data(), g.data(), g.data(), f4.data(), f4.data(), rem.data(),
                                 rem.data(), m->IareaT.data(), m->IdxCu.data(), m->IdyCv.data(),
                                 m->c.periodic_x, m->c.periodic_y));
    }
    // vert_visc_init (mom6_vert_visc_serial.F90:60-112)
    RC(mom6h_vert_visc_init(&m->vv, d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, nk,
                            mCu.data(), mCv.data(), m->c.Kv, m->c.Kv_ml, m->c.Kv_extra_bbl,
                            m->c.Hmix, m->c.Hbbl));
    // hor_visc_init (mom6_hor_visc_serial.F90:160-286)
    if (!m->mixed) {
        std::vector<double> dx2h(n2), dy2h(n2), dx2q(n2), dy2q(n2), DX_dyT(n2), DY_dxT(n2);
        std::vector<double> DX_dyBu(n2), DY_dxBu(n2), I2u(n2), Iu2(n2), I2v(n2), Iv2(n2), one(n2, 1.0);
        for (size_t q = 0; q < n2; ++q) {
            dx2h[q] = m->dxT[q] * m->dxT[q];
            dy2h[q] = m->dyT[q] * m->dyT[q];
            dx2q[q] = m->dxBu[q] * m->dxBu[q];
            dy2q[q] = m->dyBu[q] * m->dyBu[q];
            DX_dyT[q] = m->dxT[q] * m->IdyT[q];
            DY_dxT[q] = m->dyT[q] * m->IdxT[q];
            DX_dyBu[q] = m->dxBu[q] * m->IdyBu[q];
            DY_dxBu[q] = m->dyBu[q] * m->IdxBu[q];
            I2u[q] = m->IdxCu[q] * m->IdxCu[q] * m->IdyCu[q];
            Iu2[q] = m->IdxCu[q] * m->IdyCu[q] * m->IdyCu[q];
            I2v[q] = m->IdxCv[q] * m->IdxCv[q] * m->IdyCv[q];
            Iv2[q] = m->IdxCv[q] * m->IdyCv[q] * m->IdyCv[q];
        }
        const double h_neglect = 1.0e-3;  // max(Angstrom_H, 1e-3) (:203)
        RC(mom6h_hor_visc_init(&m->hv, d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, nk,
                               m->c.Kh, h_neglect, DY_dxT.data(), DX_dyT.data(), DY_dxBu.data(),
                               DX_dyBu.data(), m->IdyCu.data(), m->IdxCu.data(), m->IdyCv.data(),
                               m->IdxCv.data(), m->IareaCu.data(), m->IareaCv.data(), mT.data(),
                               mBu.data(), one.data(), one.data(), dy2h.data(), dx2h.data(),
                               dy2q.data(), dx2q.data()));
        RC(mom6h_hor_visc_set_options(m->hv, m->c.laplacian, m->c.biharmonic, m->c.Ah, m->c.no_slip,
                                      I2u.data(), Iu2.data(), I2v.data(), Iv2.data()));
        RC(mom6h_hor_visc_set_closures(m->hv, m->c.smagorinsky, m->c.leith, m->c.better_bound, m->areaT.data(),
                                       m->dxBu.data(), m->dyBu.data(), m->IdxBu.data(), m->IdyBu.data()));
    }
    // uniform-metric values for the fused kernels (exact array contents)
    {
        UMetrics &g = m->ug;
        const size_t n = n2;
        bool ok = true;
        auto U = [&](const std::vector<double> &a, double &v) { ok = ok && uniform_value(a.data(), n, &v); };
        U(m->areaT, g.areaT); U(m->IareaT, g.IareaT); U(m->dxT, g.dxT); U(m->IdxT, g.IdxT);
        U(m->dyCu, g.dyCu); U(m->dy_Cu, g.dy_Cu); U(m->dxCu, g.dxCu); U(m->IdxCu, g.IdxCu); U(m->IdyCu, g.IdyCu);
        U(m->dxCv, g.dxCv); U(m->dyCv, g.dyCv); U(m->IdxCv, g.IdxCv); U(m->IdyCv, g.IdyCv);
        U(m->IareaBu, g.IareaBu); U(m->IareaCu, g.IareaCu); U(m->IareaCv, g.IareaCv);
        U(m->dyT, g.dy2h); U(m->dxBu, g.dx2q); U(m->dyBu, g.dy2q); U(m->IdyT, g.DX_dyT);
        U(m->IdxBu, g.DY_dxBu); U(m->IdyBu, g.DX_dyBu);
        if (ok) {
            // derived exactly as hor_visc_init / coriolis_init form them
            const double dyT = m->dyT[0], dxBu = m->dxBu[0], dyBu = m->dyBu[0], IdyT = m->IdyT[0];
            const double IdxBu = m->IdxBu[0], IdyBu = m->IdyBu[0];
            g.dx2h = g.dxT * g.dxT;
            g.dy2h = dyT * dyT;
            g.dx2q = dxBu * dxBu;
            g.dy2q = dyBu * dyBu;
            g.DX_dyT = g.dxT * IdyT;
            g.DY_dxT = dyT * g.IdxT;
            g.DX_dyBu = dxBu * IdyBu;
            g.DY_dxBu = dyBu * IdxBu;
            g.Area_q = (g.areaT + g.areaT) + (g.areaT + g.areaT);
            g.Idx2dyCu = g.IdxCu * g.IdxCu * g.IdyCu;
            g.Idxdy2u = g.IdxCu * g.IdyCu * g.IdyCu;
            g.Idx2dyCv = g.IdxCv * g.IdxCv * g.IdyCv;
            g.Idxdy2v = g.IdxCv * g.IdyCv * g.IdyCv;
        }
        m->uniform = ok;
        m->fused_rows = binary_mask(mT.data(), n) && binary_mask(mBu.data(), n) &&
                        rows_uniform(fq.data(), (size_t)d.n1, (size_t)d.n2);
    }
    MOM6H_REQUIRE(!m->mixed || (m->uniform && m->fused_rows),
                  "mom6h_model_prepare: the mixed-precision form needs a grid with uniform metrics, 0/1 masks "
                  "and Coriolis constant along each row");
    // periodic fill of the initial state (rk2_mpi_serial_driver.F90:191-197)
    if (!m->mixed)
        RC(fill_halos(m, {{m->u, nk}, {m->v, nk}, {m->h, nk}, {m->h0, nk}}));
    else
        RC(fill_halos(m, {{m->f.u, nk}, {m->f.v, nk}, {m->f.h, nk}, {m->f.h0, nk}}));
    RC(fill_halos(m, {{m->eta, 1}, {m->ubt, 1}, {m->vbt, 1}}));
    MOM6H_CHECK(hipStreamSynchronize(m->s0));
    m->prepared = true;
    return MOM6H_OK;
}

// Destroys the captured step graphs (a launch shape or option they hold changed)
static void drop_graphs(mom6h_model *m)
{
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    if (m->gexec_bt) (void)hipGraphExecDestroy(m->gexec_bt);
    if (m->graph_bt) (void)hipGraphDestroy(m->graph_bt);
    m->gexec = nullptr;
    m->graph = nullptr;
    m->gexec_bt = nullptr;
    m->graph_bt = nullptr;
}

// Captures what enq enqueues on stream s as a graph and instantiates it
static int capture_on(mom6h_model *m, hipStream_t s, int (*enq)(mom6h_model *), hipGraph_t *graph,
                      hipGraphExec_t *exec)
{
    MOM6H_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enq(m);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    MOM6H_CHECK(ec);
    *graph = g;
    MOM6H_CHECK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return MOM6H_OK;
}

int mom6h_model_step(mom6h_model *m, int niter)
{
    MOM6H_REQUIRE(m != nullptr, "mom6h_model_step: NULL model");
    MOM6H_REQUIRE(m->prepared, "mom6h_model_step: call mom6h_model_prepare first");
    MOM6H_REQUIRE(niter >= 0, "mom6h_model_step: niter >= 0 required");
    // The fused path consumes vert_visc's coupling coefficients on chip; the
    // module path keeps them in the control structure, as the reference does.
    RC(mom6h_vert_visc_keep_coefficients(m->vv, fused_ok(m) ? 0 : 1));
    if (m->c.coupled) {
        // the coupled form: each continuity call waits for its btstep call, so
        // the whole step runs in order on the 3D stream (one graph when captured)
        if (m->use_graph && !m->profile) {
            const int key = fused_ok(m) ? 1 : 0;
            if (m->gexec && m->graph_fused != key) drop_graphs(m);
            if (!m->gexec) {
                RC(capture_on(m, m->s0, enqueue_step, &m->graph, &m->gexec));
                m->graph_fused = key;
            }
            for (int it = 0; it < niter; ++it) MOM6H_CHECK(hipGraphLaunch(m->gexec, m->s0));
            return MOM6H_OK;
        }
        for (int it = 0; it < niter; ++it) RC(enqueue_step(m));
        return MOM6H_OK;
    }
    if (m->use_graph && !m->profile && !m->serialize) {
        // One 3D step (s0) and one barotropic step (s1), each captured as a
        // graph on its own stream and replayed as the eager path enqueues:
        // one fork, every barotropic step on s1, every 3D step on s0, one join
        // (so a barotropic step may run beside any 3D step of the call).
        const int key = fused_ok(m) ? 1 : 0;
        if (m->gexec && m->graph_fused != key) drop_graphs(m);
        if (!m->gexec) {
            RC(capture_on(m, m->s0, enqueue_step, &m->graph, &m->gexec));
            const int rc = capture_on(m, m->s1, enqueue_barotropic, &m->graph_bt, &m->gexec_bt);
            if (rc) {
                drop_graphs(m);
                return rc;
            }
            m->graph_fused = key;
        }
        MOM6H_CHECK(hipEventRecord(m->e_start, m->s0));
        MOM6H_CHECK(hipStreamWaitEvent(m->s1, m->e_start, 0));
        for (int it = 0; it < niter; ++it) {
            MOM6H_CHECK(hipGraphLaunch(m->gexec_bt, m->s1));
            MOM6H_CHECK(hipGraphLaunch(m->gexec, m->s0));
        }
        MOM6H_CHECK(hipEventRecord(m->e_bt, m->s1));
        MOM6H_CHECK(hipStreamWaitEvent(m->s0, m->e_bt, 0));
        return MOM6H_OK;
    }
    if (m->serialize) {
        for (int it = 0; it < niter; ++it) {
            RC(enqueue_barotropic(m));
            RC(enqueue_step(m));
        }
        return MOM6H_OK;
    }
    // The barotropic chain and the 3D chain share no array (the 2D fields the
    // barotropic solver reads and writes are never touched by the 3D stages),
    // so s1 runs all its steps without waiting for s0; the streams join once.
    // The steps are enqueued in turn, a barotropic step then a 3D step: the
    // host takes about as long to enqueue a barotropic step (some thirty
    // launches) as the GPU to run it, so enqueueing all barotropic steps
    // first would leave s0 idle until the last was enqueued, and the two
    // chains would run one after the other instead of side by side.
    MOM6H_CHECK(hipEventRecord(m->e_start, m->s0));
    MOM6H_CHECK(hipStreamWaitEvent(m->s1, m->e_start, 0));
    for (int it = 0; it < niter; ++it) {
        RC(enqueue_barotropic(m));
        RC(enqueue_step(m));
    }
    MOM6H_CHECK(hipEventRecord(m->e_bt, m->s1));
    MOM6H_CHECK(hipStreamWaitEvent(m->s0, m->e_bt, 0));
    return MOM6H_OK;
}

int mom6h_model_synchronize(mom6h_model *m)
{
    MOM6H_REQUIRE(m != nullptr, "mom6h_model_synchronize: NULL model");
    MOM6H_CHECK(hipStreamSynchronize(m->s1));
    MOM6H_CHECK(hipStreamSynchronize(m->s0));
    if (m->profile && !m->ev.empty()) {
        for (size_t q = 0; q + 2 < m->ev.size(); q += 3) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, m->ev[q], m->ev[q + 1]) == hipSuccess)
                m->stage_ms[(int)(intptr_t)m->ev[q + 2]] += ms;
            (void)hipEventDestroy(m->ev[q]);
            (void)hipEventDestroy(m->ev[q + 1]);
        }
        m->ev.clear();
    }
    return MOM6H_OK;
}

int mom6h_model_set_option(mom6h_model *m, const char *key, double value)
{
    MOM6H_REQUIRE(m && key, "mom6h_model_set_option: NULL argument");
    const std::string k(key);
    // Steps already enqueued read the launch shape and may be replaying a
    // graph about to be destroyed: let them finish before anything changes.
    if (m->s0) MOM6H_CHECK(hipStreamSynchronize(m->s0));
    if (m->s1) MOM6H_CHECK(hipStreamSynchronize(m->s1));
    if (k == "graph") m->use_graph = value != 0.0;
    else if (k == "serialize") m->serialize = value != 0.0;
    else if (k == "bt_tile") {
        MOM6H_REQUIRE(m->bt_cs, "mom6h_model_set_option: bt_tile after prepare");
        MOM6H_REQUIRE(value == 4.0 || value == 8.0 || value == 16.0, "mom6h_model_set_option: bt_tile must be 4, 8 or 16");
        m->bt_cs->tile_rows = (int)value;
        drop_graphs(m);
    }
    else if (k == "bt_rows") {
        MOM6H_REQUIRE(m->bt_cs, "mom6h_model_set_option: bt_rows after prepare");
        m->bt_cs->use_rows = value != 0.0;
        drop_graphs(m);
    }
    else if (k == "bt_substeps") {
        MOM6H_REQUIRE(m->bt_cs, "mom6h_model_set_option: bt_substeps after prepare");
        MOM6H_REQUIRE(value >= 1.0 && value <= 5.0 && value == (double)(int)value,
                      "mom6h_model_set_option: bt_substeps must be 1..5");
        m->bt_cs->substeps = (int)value;
        drop_graphs(m);
    }
    else if (k == "momentum_tile" || k == "momentum_splits") {
        const int v = (int)value;
        if (k == "momentum_tile") {
            MOM6H_REQUIRE(v == 4 || v == 8 || v == 16, "mom6h_model_set_option: momentum_tile must be 4, 8 or 16");
            m->momentum_tile = v;
        } else {
            MOM6H_REQUIRE(v >= 1 && v <= m->c.nk, "mom6h_model_set_option: momentum_splits must be 1..nk");
            m->momentum_splits = v;
        }
        drop_graphs(m);   // the captured step holds the old launch shape
    }
    else if (k == "general") {   // module hor_visc and CorAdCalc: the general kernels on any grid
        if (m->hv) m->hv->general = value != 0.0;
        if (m->cor) m->cor->general = value != 0.0;
        drop_graphs(m);
    }
    else if (k == "fused") {
        MOM6H_REQUIRE(!m->mixed || value != 0.0, "mom6h_model_set_option: the mixed-precision form is fused only");
        m->fused = value != 0.0;
    }
    else if (k == "profile") {
        m->profile = value != 0.0;
        for (double &x : m->stage_ms) x = 0.0;
    }
    else if (k == "diag") {
        MOM6H_REQUIRE(!m->mixed || value == 0.0, "mom6h_model_set_option: diagnostics need the FP64 form");
        if (value != 0.0 && !m->diag_out) {
            const size_t n2 = (size_t)m->d.s2, n3 = n2 * m->c.nk;
            RC(dev_alloc(&m->diag_work2d, n2, true));   // the reference's work_2d starts at zero
            RC(dev_alloc(&m->diag_ke, n3, true));
            RC(dev_alloc(&m->diag_mass, n2, true));
            RC(dev_alloc(&m->diag_red, (size_t)mom6h_diag_stats_work_doubles(), true));
            RC(dev_alloc(&m->diag_out, 16, true));
            m->fields.push_back(Field{"diag_work2d", (void **)&m->diag_work2d, n2, 8, false});
            m->fields.push_back(Field{"diag_ke", (void **)&m->diag_ke, n3, 8, false});
            m->fields.push_back(Field{"diag_mass", (void **)&m->diag_mass, n2, 8, false});
        }
        m->diag = value != 0.0;
        drop_graphs(m);   // the captured step holds the posts or not
    } else return set_error(MOM6H_EINVAL, "mom6h_model_set_option: unknown option '%s'", key);
    return MOM6H_OK;
}

int mom6h_model_diagnostics(mom6h_model *m, double *out)
{
    MOM6H_REQUIRE(m && out, "mom6h_model_diagnostics: NULL argument");
    MOM6H_REQUIRE(m->diag && m->prepared, "mom6h_model_diagnostics: set option diag and prepare first");
    const Dims &d = m->d;
    hipStream_t s0 = m->s0;
    MOM6H_CHECK(hipStreamSynchronize(m->s1));
    // end-of-run diagnostics (rk2_serial_driver.F90:386-394): KE, then mass
    RC(mom6h_diag_ke(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, d.nk, m->u, m->v, m->diag_ke, s0));
    RC(mom6h_diag_stats(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, d.nk, m->diag_ke, m->diag_red,
                        m->diag_out, s0));
    RC(mom6h_diag_mass(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, d.nk, m->h, m->diag_mass, s0));
    RC(mom6h_diag_stats(d.isd, d.ied, d.jsd, d.jed, d.isc, d.iec, d.jsc, d.jec, 1, m->diag_mass, m->diag_red,
                        m->diag_out + 12, s0));
    MOM6H_CHECK(hipMemcpyAsync(out, m->diag_out, 16 * sizeof(double), hipMemcpyDeviceToHost, s0));
    MOM6H_CHECK(hipStreamSynchronize(s0));
    return MOM6H_OK;
}

int mom6h_model_stage_times(mom6h_model *m, double *ms, int max_stages, char *names, int nbytes)
{
    MOM6H_REQUIRE(m != nullptr, "mom6h_model_stage_times: NULL model");
    std::string all;
    for (int s = 0; s < ST_COUNT; ++s) {
        if (ms && s < max_stages) ms[s] = m->stage_ms[s];
        if (s) all += ",";
        all += kStageNames[s];
    }
    if (names && nbytes > 0) {
        strncpy(names, all.c_str(), (size_t)nbytes - 1);
        names[nbytes - 1] = 0;
    }
    return ST_COUNT;
}

int mom6h_model_main_stream(mom6h_model *m, void **stream)
{
    MOM6H_REQUIRE(m && stream, "mom6h_model_main_stream: NULL argument");
    *stream = (void *)m->s0;
    return MOM6H_OK;
}

int mom6h_model_autotune(mom6h_model *m, int reps, char *report, int nbytes)
{
    MOM6H_REQUIRE(m != nullptr, "mom6h_model_autotune: NULL model");
    MOM6H_REQUIRE(m->prepared, "mom6h_model_autotune: call mom6h_model_prepare first");
    MOM6H_REQUIRE(fused_ok(m), "mom6h_model_autotune: only the fused stencil has tile shapes to choose");
    MOM6H_CHECK(hipDeviceSynchronize());
    const int saved_profile = m->profile;
    m->profile = 0;
    // the predictor's momentum stencil (u, v, h0 -> up, vp): every step rewrites its outputs
    auto launch = [&]() -> int {
        return m->mixed ? momentum_phase_f32(m, m->f.u, m->f.v, m->f.h0, m->f.up, m->f.vp, m->c.dt)
                        : momentum_phase(m, m->u, m->v, m->h0, m->up, m->vp, m->c.dt);
    };
    hipEvent_t a, b;
    MOM6H_CHECK(hipEventCreate(&a));
    MOM6H_CHECK(hipEventCreate(&b));
    const int rows[3] = {4, 8, 16}, splits[4] = {1, 2, 4, 8};
    float best = 0.f;
    int pick = 8, pick_s = 1, rc = 0;
    std::string rep = "momentum tile rows x layer splits (ms):";
    for (int q = 0; q < 12 && !rc; ++q) {
        m->momentum_tile = rows[q / 4];
        m->momentum_splits = splits[q % 4] < m->c.nk ? splits[q % 4] : m->c.nk;
        rc = launch();   // warm
        if (!rc && hipEventRecord(a, m->s0) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event record failed");
        for (int r = 0; r < reps && !rc; ++r) rc = launch();
        if (!rc && hipEventRecord(b, m->s0) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event record failed");
        if (!rc && hipEventSynchronize(b) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event wait failed");
        float ms = 0.f;
        if (!rc && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event timing failed");
        ms /= (float)(reps > 0 ? reps : 1);
        char buf[64];
        snprintf(buf, sizeof(buf), " %dx%d %.4f;", m->momentum_tile, m->momentum_splits, ms);
        rep += buf;
        if (q == 0 || ms < best) {
            best = ms;
            pick = m->momentum_tile;
            pick_s = m->momentum_splits;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    m->profile = saved_profile;
    m->momentum_tile = pick;
    m->momentum_splits = pick_s;
    drop_graphs(m);
    if (rc) return rc;
    rep += " chosen " + std::to_string(pick) + "x" + std::to_string(pick_s);
    if (report && nbytes > 0) snprintf(report, (size_t)nbytes, "%s", rep.c_str());
    return MOM6H_OK;
}

int mom6h_model_destroy(mom6h_model *m)
{
    if (!m) return MOM6H_OK;
    if (m->s0) (void)hipStreamSynchronize(m->s0);
    if (m->s1) (void)hipStreamSynchronize(m->s1);
    mom6h_continuity_end(m->cont);
    mom6h_coriolis_end(m->cor);
    mom6h_barotropic_end(m->bt_cs);
    mom6h_vert_visc_end(m->vv);
    mom6h_hor_visc_end(m->hv);
    for (auto &f : m->fields)
        if (f.owned && *f.p) (void)hipFree(*f.p);
    mom6h_bt_cont_free(&m->bt);
    mom6h_bt_cont_free(&m->bt_v);
    for (double *p : {m->diag_work2d, m->diag_ke, m->diag_mass, m->diag_red, m->diag_out}) dev_free(p);
    drop_graphs(m);
    if (m->e_start) (void)hipEventDestroy(m->e_start);
    if (m->e_bt) (void)hipEventDestroy(m->e_bt);
    if (m->s0) (void)hipStreamDestroy(m->s0);
    if (m->s1) (void)hipStreamDestroy(m->s1);
    delete m;
    return MOM6H_OK;
}

}  // extern "C"

namespace mom6h {
bool fused_ok(const mom6h_model *m)
{
    return m->fused && m->uniform && m->fused_rows && m->c.laplacian && m->c.coriolis_scheme == 1 &&
           !m->c.smagorinsky && !m->c.leith && !m->c.better_bound;
}
}  // namespace mom6h
