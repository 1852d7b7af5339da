// mom6h_ensemble.hip -- an ensemble of split-RK2 runs advanced together.
//
// Members are whole models (mom6h_model: own grid size, nk, state, control
// structures) that share bt_nsteps, the precision form and the switches that
// select kernels (meridional pass, biharmonic viscosity).  A step runs the
// reference driver's stage order (app/rk2_serial_driver.F90:218-399) once for
// all members: every stage is ONE launch whose blocks cover the tiles of
// every member (BatchMap), each block working on its member with that
// member's argument block, read from a device table built at prepare time.
// The barotropic chain of all members runs on a second stream, as in the
// single model.  Each member gets the arithmetic of its single-model step, so
// batched and single results agree bit for bit.
#include <cstdio>
#include <cstring>

#include "mom6h_model.h"

using namespace mom6h;

namespace {

// a host table mirrored on the device
template <typename A>
struct Table {
    std::vector<A> h;
    A *d = nullptr;
    int upload()
    {
        if (h.empty()) return MOM6H_OK;
        MOM6H_CHECK(hipMalloc((void **)&d, h.size() * sizeof(A)));
        MOM6H_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(A), hipMemcpyHostToDevice));
        return MOM6H_OK;
    }
    void release()
    {
        if (d) (void)hipFree(d);
        d = nullptr;
    }
    int n() const { return (int)h.size(); }
};

// the layer fields of one member in the ensemble's arithmetic type
template <typename T>
struct Layers {
    T *u, *v, *h, *h0, *htmp, *up, *vp, *uh, *vh, *u_cor, *bt_h_u, *hx, *v_cor, *bt_h_v, *del2u, *del2v;
};

Layers<float> layers_f32(mom6h_model *m)
{
    auto &f = m->f;
    return {f.u, f.v, f.h, f.h0, f.htmp, f.up, f.vp, f.uh, f.vh, f.u_cor, f.bt_h_u, f.hx, f.v_cor, f.bt_h_v,
            f.del2u, f.del2v};
}

Layers<double> layers_f64(mom6h_model *m)
{
    return {m->u, m->v, m->h, m->h0, m->htmp, m->up, m->vp, m->uh, m->vh, m->u_cor, m->bt.h_u, m->hx, m->v_cor,
            m->bt_v.h_u, m->del2u, m->del2v};
}

template <typename T>
struct StepTables {
    Table<MomArgsT<T>> mom_pred, mom_corr;
    Table<VVArgs<T>> vv_pred, vv_corr;
    Table<ContArgsT<T>> cx_pred, cx_corr, cy_pred, cy_corr;
    Table<HaloJob<T>> h_upvp, h_htmp, h_uvh, h_hx;
    void release()
    {
        for (auto *t : {&mom_pred, &mom_corr}) t->release();
        for (auto *t : {&vv_pred, &vv_corr}) t->release();
        for (auto *t : {&cx_pred, &cx_corr, &cy_pred, &cy_corr}) t->release();
        for (auto *t : {&h_upvp, &h_htmp, &h_uvh, &h_hx}) t->release();
    }
};

}  // namespace

struct mom6h_ensemble {
    std::vector<mom6h_model *> members;
    bool prepared = false, mixed = false, meridional = false;
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t e_start = nullptr, e_bt = nullptr;
    int use_graph = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;      // one 3D step of every member (stream s0)
    hipGraph_t graph_bt = nullptr;       // one barotropic step of every member (stream s1)
    hipGraphExec_t gexec_bt = nullptr;
    int momentum_tile = 8;
    int bt_tile = 4;         // rows of the barotropic substep tile (4, 8 or 16)
    int bt_substeps = 5;     // barotropic substeps per blocked launch (1 = one launch per substep)
    long long max_lds = 0;
    StepTables<float> tf;
    StepTables<double> td;
    Table<BTMember> bt_pred, bt_corr;
    Table<HaloJob<double>> h2d;
};

namespace {

#define RC(expr)               \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

template <typename T>
HaloJob<T> job(const mom6h_model *m, T *a, int nk)
{
    return HaloJob<T>{m->d, a, nk, m->halo.px, m->halo.py};
}

template <typename T>
int build_tables(mom6h_ensemble *e, StepTables<T> &t, Layers<T> (*layers)(mom6h_model *))
{
    for (mom6h_model *m : e->members) {
        const Layers<T> L = layers(m);
        const double dt = m->c.dt;
        const int nk = m->c.nk;
        MomArgsT<T> a{};
        a.d = m->d; a.g = m->ug; a.Kh = (T)m->c.Kh; a.h_neglect = (T)1.0e-3; a.no_slip = m->c.no_slip;
        a.mask2dT = m->mask2dT; a.mask2dBu = m->mask2dBu; a.CoriolisBu = m->CoriolisBu;
        a.bu = L.u; a.bv = L.v;
        a.Ah = (T)m->c.Ah; a.del2u = m->c.biharmonic ? L.del2u : nullptr; a.del2v = m->c.biharmonic ? L.del2v : nullptr;
        a.nkz = 1;
        // predictor: u, v, h0 -> up, vp; corrector: up, vp, htmp -> u, v (rk2_serial_driver.F90:236-350)
        a.u = L.u; a.v = L.v; a.h = L.h0; a.ou = L.up; a.ov = L.vp; a.fdt = (T)dt;
        t.mom_pred.h.push_back(a);
        a.u = L.up; a.v = L.vp; a.h = L.htmp; a.ou = L.u; a.ov = L.v; a.fdt = (T)(0.5 * dt);
        t.mom_corr.h.push_back(a);
        const mom6h_vert_visc_cs *vv = m->vv;
        VVArgs<T> p{m->d, L.up, L.vp, L.h0, dt, m->taux, m->tauy, m->ray_u, m->ray_v, vv->mask2dCu, vv->mask2dCv,
                    vv->a_u, vv->a_v, vv->h_u, vv->h_v, vv->taux_bot, vv->tauy_bot, vv->Kv, vv->Kv_ml,
                    vv->Kv_extra_bbl, vv->Hmix, vv->Hbbl};
        t.vv_pred.h.push_back(p);
        p.u = L.u; p.v = L.v; p.h = L.htmp; p.dt = 0.5 * dt;
        t.vv_corr.h.push_back(p);
        // continuity: with the meridional pass the zonal one writes hx
        const mom6h_continuity_cs *cs = m->cont;
        T *x_out_p = e->meridional ? L.hx : L.htmp, *x_out_c = e->meridional ? L.hx : L.h;
        t.cx_pred.h.push_back(continuity_args<T>(cs, false, L.up, L.h0, x_out_p, L.uh, dt, &m->bt, L.bt_h_u, L.u_cor,
                                                 m->du_cor));
        t.cx_corr.h.push_back(continuity_args<T>(cs, false, L.u, L.htmp, x_out_c, L.uh, 0.5 * dt, &m->bt, L.bt_h_u,
                                                 L.u_cor, m->du_cor));
        if (e->meridional) {
            t.cy_pred.h.push_back(continuity_args<T>(cs, true, L.vp, L.hx, L.htmp, L.vh, dt, &m->bt_v, L.bt_h_v,
                                                     L.v_cor, m->dv_cor));
            t.cy_corr.h.push_back(continuity_args<T>(cs, true, L.v, L.hx, L.h, L.vh, 0.5 * dt, &m->bt_v, L.bt_h_v,
                                                     L.v_cor, m->dv_cor));
            t.h_hx.h.push_back(job(m, L.hx, nk));
        }
        t.h_upvp.h.push_back(job(m, L.up, nk));
        t.h_upvp.h.push_back(job(m, L.vp, nk));
        t.h_htmp.h.push_back(job(m, L.htmp, nk));
        for (T *x : {L.u, L.v, L.h}) t.h_uvh.h.push_back(job(m, x, nk));
    }
    for (auto *x : {&t.mom_pred, &t.mom_corr}) RC(x->upload());
    for (auto *x : {&t.vv_pred, &t.vv_corr}) RC(x->upload());
    for (auto *x : {&t.cx_pred, &t.cx_corr, &t.cy_pred, &t.cy_corr}) RC(x->upload());
    for (auto *x : {&t.h_upvp, &t.h_htmp, &t.h_uvh, &t.h_hx}) RC(x->upload());
    return MOM6H_OK;
}

template <typename T>
int halo(const Table<HaloJob<T>> &t, hipStream_t s)
{
    return t.n() ? halo_batch<T>(t.d, t.h.data(), t.n(), s) : MOM6H_OK;
}

// the 3D chain of one step for all members (stream s0)
template <typename T>
int enqueue_3d(mom6h_ensemble *e, StepTables<T> &t)
{
    hipStream_t s = e->s0;
    const int n = (int)e->members.size();
    const int ty = e->momentum_tile;
    // predictor
    RC(launch_momentum_batch<T>(t.mom_pred.d, t.mom_pred.h.data(), n, s, ty));
    RC(vert_visc_batch<T>(t.vv_pred.d, t.vv_pred.h.data(), n, e->max_lds, s));
    RC(halo(t.h_upvp, s));
    RC(continuity_batch<T>(t.cx_pred.d, t.cx_pred.h.data(), n, false, true, e->max_lds, s));
    if (e->meridional) {
        RC(halo(t.h_hx, s));
        RC(continuity_batch<T>(t.cy_pred.d, t.cy_pred.h.data(), n, true, true, e->max_lds, s));
    }
    RC(halo(t.h_htmp, s));
    // corrector
    RC(launch_momentum_batch<T>(t.mom_corr.d, t.mom_corr.h.data(), n, s, ty));
    RC(vert_visc_batch<T>(t.vv_corr.d, t.vv_corr.h.data(), n, e->max_lds, s));
    RC(continuity_batch<T>(t.cx_corr.d, t.cx_corr.h.data(), n, false, true, e->max_lds, s));
    if (e->meridional) {
        RC(halo(t.h_hx, s));
        RC(continuity_batch<T>(t.cy_corr.d, t.cy_corr.h.data(), n, true, true, e->max_lds, s));
    }
    RC(halo(t.h_uvh, s));
    return MOM6H_OK;
}

// the barotropic chain of one step for all members (stream s1)
int enqueue_bt(mom6h_ensemble *e)
{
    const int n = (int)e->members.size();
    RC(btstep_batch(e->bt_pred.d, e->bt_pred.h.data(), n, e->s1, e->bt_tile, e->bt_substeps));
    RC(btstep_batch(e->bt_corr.d, e->bt_corr.h.data(), n, e->s1, e->bt_tile, e->bt_substeps));
    return halo(e->h2d, e->s1);
}

// one 3D step of every member on s0
int enqueue_3d_step(mom6h_ensemble *e)
{
    return e->mixed ? enqueue_3d<float>(e, e->tf) : enqueue_3d<double>(e, e->td);
}

// Captures what enq enqueues on stream s as a graph and instantiates it
int capture_on(mom6h_ensemble *e, hipStream_t s, int (*enq)(mom6h_ensemble *), hipGraph_t *graph,
               hipGraphExec_t *exec)
{
    MOM6H_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enq(e);
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

// layer splits of every member's momentum launch (capped at its nk), both phases
template <typename T>
int set_splits(StepTables<T> &t, int nkz)
{
    for (auto *x : {&t.mom_pred, &t.mom_corr}) {
        for (auto &a : x->h) a.nkz = nkz < a.d.nk ? nkz : a.d.nk;
        if (x->d) MOM6H_CHECK(hipMemcpy(x->d, x->h.data(), x->h.size() * sizeof(x->h[0]), hipMemcpyHostToDevice));
    }
    return MOM6H_OK;
}

void drop_graph(mom6h_ensemble *e)
{
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->gexec_bt) (void)hipGraphExecDestroy(e->gexec_bt);
    if (e->graph_bt) (void)hipGraphDestroy(e->graph_bt);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->gexec_bt = nullptr;
    e->graph_bt = nullptr;
}

// mean time (ms) of reps predictor momentum launches with tile height ty
template <typename T>
int time_momentum(mom6h_ensemble *e, StepTables<T> &t, int ty, int reps, float *ms)
{
    const int n = (int)e->members.size();
    hipEvent_t a, b;
    MOM6H_CHECK(hipEventCreate(&a));
    MOM6H_CHECK(hipEventCreate(&b));
    int rc = launch_momentum_batch<T>(t.mom_pred.d, t.mom_pred.h.data(), n, e->s0, ty);   // warm
    if (!rc && hipEventRecord(a, e->s0) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event record failed");
    for (int r = 0; r < reps && !rc; ++r) rc = launch_momentum_batch<T>(t.mom_pred.d, t.mom_pred.h.data(), n, e->s0, ty);
    if (!rc && hipEventRecord(b, e->s0) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event record failed");
    if (!rc && hipEventSynchronize(b) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event wait failed");
    float x = 0.f;
    if (!rc && hipEventElapsedTime(&x, a, b) != hipSuccess) rc = set_error(MOM6H_ESTATE, "event timing failed");
    *ms = x / (float)(reps > 0 ? reps : 1);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace

extern "C" {

int mom6h_ensemble_create(mom6h_ensemble **out, const mom6h_config *member_cfg, int n)
{
    MOM6H_REQUIRE(out && member_cfg, "mom6h_ensemble_create: NULL argument");
    MOM6H_REQUIRE(n >= 1 && n <= 21, "mom6h_ensemble_create: 1..21 members (3 halo jobs each in one launch)");
    const mom6h_config &c0 = member_cfg[0];
    for (int q = 0; q < n; ++q)
        MOM6H_REQUIRE(!member_cfg[q].smagorinsky && !member_cfg[q].leith && !member_cfg[q].better_bound &&
                          !member_cfg[q].coupled && !member_cfg[q].ray3d,
                      "mom6h_ensemble_create: the batched stencils have no Smagorinsky/Leith/bounded viscosity "
                      "and no coupled form or 3D Rayleigh drag");
    for (int q = 1; q < n; ++q) {
        const mom6h_config &c = member_cfg[q];
        MOM6H_REQUIRE(c.bt_nsteps == c0.bt_nsteps && c.precision == c0.precision && c.meridional == c0.meridional &&
                          c.biharmonic == c0.biharmonic,
                      "mom6h_ensemble_create: members must share bt_nsteps, precision, meridional and biharmonic");
    }
    auto *e = new mom6h_ensemble();
    e->mixed = c0.precision == 1;
    e->meridional = c0.meridional != 0;
    int rc = 0;
    for (int q = 0; q < n && !rc; ++q) {
        mom6h_model *m = nullptr;
        rc = mom6h_model_create(&m, &member_cfg[q]);
        if (!rc) e->members.push_back(m);
    }
    if (!rc && hipStreamCreateWithFlags(&e->s0, hipStreamNonBlocking) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "stream creation failed");
    if (!rc && hipStreamCreateWithFlags(&e->s1, hipStreamNonBlocking) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "stream creation failed");
    if (!rc && hipEventCreateWithFlags(&e->e_start, hipEventDisableTiming) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "event creation failed");
    if (!rc && hipEventCreateWithFlags(&e->e_bt, hipEventDisableTiming) != hipSuccess)
        rc = set_error(MOM6H_ENOMEM, "event creation failed");
    if (rc) {
        mom6h_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return MOM6H_OK;
}

int mom6h_ensemble_member(mom6h_ensemble *e, int index, mom6h_model **m)
{
    MOM6H_REQUIRE(e && m, "mom6h_ensemble_member: NULL argument");
    MOM6H_REQUIRE(index >= 0 && index < (int)e->members.size(), "mom6h_ensemble_member: index %d out of range",
                  index);
    *m = e->members[index];
    return MOM6H_OK;
}

int mom6h_ensemble_prepare(mom6h_ensemble *e)
{
    MOM6H_REQUIRE(e != nullptr, "mom6h_ensemble_prepare: NULL ensemble");
    MOM6H_REQUIRE(!e->prepared, "mom6h_ensemble_prepare: already prepared");
    for (mom6h_model *m : e->members) {
        RC(mom6h_model_prepare(m));
        MOM6H_REQUIRE(fused_ok(m), "mom6h_ensemble_prepare: every member needs the fused stencils (uniform metrics, "
                                   "0/1 masks, Coriolis constant along rows, Sadourny Coriolis, Laplacian viscosity)");
        MOM6H_REQUIRE(m->bt_cs->uniform, "mom6h_ensemble_prepare: members need uniform barotropic coefficients");
        e->max_lds = m->vv->max_lds;
    }
    RC(e->mixed ? build_tables<float>(e, e->tf, layers_f32) : build_tables<double>(e, e->td, layers_f64));
    for (mom6h_model *m : e->members) {
        BTMember p{}, c{};
        RC(bt_member(m->bt_cs, m->eta, m->ubt, m->vbt, m->ubt_av, m->vbt_av, m->eta_av, &p));
        RC(bt_member(m->bt_cs, m->eta_av, m->ubt_av, m->vbt_av, m->ubt_av, m->vbt_av, m->eta, &c));
        e->bt_pred.h.push_back(p);
        e->bt_corr.h.push_back(c);
        for (double *x : {m->eta, m->ubt, m->vbt}) e->h2d.h.push_back(job(m, x, 1));
    }
    RC(e->bt_pred.upload());
    RC(e->bt_corr.upload());
    RC(e->h2d.upload());
    MOM6H_CHECK(hipDeviceSynchronize());
    e->prepared = true;
    return MOM6H_OK;
}

int mom6h_ensemble_step(mom6h_ensemble *e, int niter)
{
    MOM6H_REQUIRE(e != nullptr, "mom6h_ensemble_step: NULL ensemble");
    MOM6H_REQUIRE(e->prepared, "mom6h_ensemble_step: call mom6h_ensemble_prepare first");
    MOM6H_REQUIRE(niter >= 0, "mom6h_ensemble_step: niter >= 0 required");
    // The barotropic chain and the 3D chain share no array: one fork, every
    // barotropic step on s1, every 3D step on s0, one join (as one graph per
    // stream and step with the graph option)
    if (e->use_graph && !e->gexec) {
        RC(capture_on(e, e->s0, enqueue_3d_step, &e->graph, &e->gexec));
        const int rc = capture_on(e, e->s1, enqueue_bt, &e->graph_bt, &e->gexec_bt);
        if (rc) {
            drop_graph(e);
            return rc;
        }
    }
    MOM6H_CHECK(hipEventRecord(e->e_start, e->s0));
    MOM6H_CHECK(hipStreamWaitEvent(e->s1, e->e_start, 0));
    // enqueued in turn, so that s0 has work while the host still enqueues s1's
    for (int it = 0; it < niter; ++it) {
        if (e->use_graph) MOM6H_CHECK(hipGraphLaunch(e->gexec_bt, e->s1));
        else RC(enqueue_bt(e));
        if (e->use_graph) MOM6H_CHECK(hipGraphLaunch(e->gexec, e->s0));
        else RC(enqueue_3d_step(e));
    }
    MOM6H_CHECK(hipEventRecord(e->e_bt, e->s1));
    MOM6H_CHECK(hipStreamWaitEvent(e->s0, e->e_bt, 0));
    return MOM6H_OK;
}

int mom6h_ensemble_synchronize(mom6h_ensemble *e)
{
    MOM6H_REQUIRE(e != nullptr, "mom6h_ensemble_synchronize: NULL ensemble");
    MOM6H_CHECK(hipStreamSynchronize(e->s1));
    MOM6H_CHECK(hipStreamSynchronize(e->s0));
    return MOM6H_OK;
}

int mom6h_ensemble_set_option(mom6h_ensemble *e, const char *key, double value)
{
    MOM6H_REQUIRE(e && key, "mom6h_ensemble_set_option: NULL argument");
    const std::string k(key);
    // Steps already enqueued read the argument tables rewritten below and may
    // be replaying a graph about to be destroyed: let them finish first.
    if (e->s0) MOM6H_CHECK(hipStreamSynchronize(e->s0));
    if (e->s1) MOM6H_CHECK(hipStreamSynchronize(e->s1));
    if (k == "graph") {
        e->use_graph = value != 0.0;
    } else if (k == "momentum_tile") {
        const int ty = (int)value;
        MOM6H_REQUIRE(ty == 4 || ty == 8 || ty == 16, "mom6h_ensemble_set_option: momentum_tile must be 4, 8 or 16");
        e->momentum_tile = ty;
        drop_graph(e);
    } else if (k == "bt_tile") {
        MOM6H_REQUIRE(value == 4.0 || value == 8.0 || value == 16.0, "mom6h_ensemble_set_option: bt_tile must be 4, 8 or 16");
        e->bt_tile = (int)value;
        drop_graph(e);
    } else if (k == "bt_substeps") {
        MOM6H_REQUIRE(value >= 1.0 && value <= 5.0 && value == (double)(int)value,
                      "mom6h_ensemble_set_option: bt_substeps must be 1..5");
        e->bt_substeps = (int)value;
        drop_graph(e);
    } else if (k == "momentum_splits") {
        MOM6H_REQUIRE(e->prepared, "mom6h_ensemble_set_option: momentum_splits after prepare");
        const int v = (int)value;
        MOM6H_REQUIRE(v >= 1, "mom6h_ensemble_set_option: momentum_splits >= 1");
        RC(e->mixed ? set_splits<float>(e->tf, v) : set_splits<double>(e->td, v));
        drop_graph(e);
    } else {
        return set_error(MOM6H_EINVAL, "mom6h_ensemble_set_option: unknown option '%s'", key);
    }
    return MOM6H_OK;
}

int mom6h_ensemble_autotune(mom6h_ensemble *e, int reps, char *report, int nbytes)
{
    MOM6H_REQUIRE(e != nullptr, "mom6h_ensemble_autotune: NULL ensemble");
    MOM6H_REQUIRE(e->prepared, "mom6h_ensemble_autotune: call mom6h_ensemble_prepare first");
    MOM6H_CHECK(hipDeviceSynchronize());
    const int rows[3] = {4, 8, 16}, splits[4] = {1, 2, 4, 8};
    float best = 0.f;
    int pick = 8, pick_s = 1;
    std::string rep = "momentum tile rows x layer splits (ms):";
    for (int q = 0; q < 12; ++q) {
        float ms = 0.f;
        RC(e->mixed ? set_splits<float>(e->tf, splits[q % 4]) : set_splits<double>(e->td, splits[q % 4]));
        RC(e->mixed ? time_momentum<float>(e, e->tf, rows[q / 4], reps, &ms)
                    : time_momentum<double>(e, e->td, rows[q / 4], reps, &ms));
        char buf[64];
        snprintf(buf, sizeof(buf), " %dx%d %.4f;", rows[q / 4], splits[q % 4], ms);
        rep += buf;
        if (q == 0 || ms < best) {
            best = ms;
            pick = rows[q / 4];
            pick_s = splits[q % 4];
        }
    }
    e->momentum_tile = pick;
    RC(e->mixed ? set_splits<float>(e->tf, pick_s) : set_splits<double>(e->td, pick_s));
    drop_graph(e);
    rep += " chosen " + std::to_string(pick) + "x" + std::to_string(pick_s);
    if (report && nbytes > 0) snprintf(report, (size_t)nbytes, "%s", rep.c_str());
    return MOM6H_OK;
}

int mom6h_ensemble_destroy(mom6h_ensemble *e)
{
    if (!e) return MOM6H_OK;
    if (e->s0) (void)hipStreamSynchronize(e->s0);
    if (e->s1) (void)hipStreamSynchronize(e->s1);
    drop_graph(e);
    e->tf.release();
    e->td.release();
    e->bt_pred.release();
    e->bt_corr.release();
    e->h2d.release();
    for (mom6h_model *m : e->members) mom6h_model_destroy(m);
    if (e->e_start) (void)hipEventDestroy(e->e_start);
    if (e->e_bt) (void)hipEventDestroy(e->e_bt);
    if (e->s0) (void)hipStreamDestroy(e->s0);
    if (e->s1) (void)hipStreamDestroy(e->s1);
    delete e;
    return MOM6H_OK;
}

}  // extern "C"
