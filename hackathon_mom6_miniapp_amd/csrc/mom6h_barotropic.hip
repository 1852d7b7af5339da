// mom6h_barotropic.hip -- the barotropic substep loop as ONE persistent kernel.
//
// Arithmetic of src/serial/mom6_barotropic_serial.F90:178-356 (btstep,
// update_u, update_v) and, for periodic grids, the split API's extended
// transport ranges (btstep_do_step, :379-474) with a refill of the ubt, vbt and
// eta halos after every substep but the last (app/rk2_mpi_serial_driver.F90:
// 297-306).  The reference CUDA-C path issues ~9 launches per substep from the
// host; here a cooperative launch runs all substeps with grid-wide barriers
// between the dependent phases of one substep:
//   P1  store previous velocities, eta predictor
//   P2  pressure forces, first velocity update (v on odd substeps)
//   P3  second velocity update, transports uhbt/vhbt
//   P4  free-surface update, time averages
//   P5  (periodic only) halo refill
// Every phase writes only cells that no other thread of the same phase reads.
// State/work arrays persist in the control structure and start at zero, so
// the cells the reference never writes (uhbt(is-1), uhbt(ie), vhbt(js-1),
// vhbt(je) with the narrow ranges) read as zero, as they do there.
#include <hip/hip_cooperative_groups.h>

#include "mom6h_internal.h"

namespace cg = cooperative_groups;

namespace mom6h {

struct BT {
    Dims d;
    int nstep, first_direction, extended, px, py;
    double dtbt, bebt, dgeo_de;
    const double *Datu, *Datv, *gtot_E, *gtot_W, *gtot_N, *gtot_S, *f_4_u, *f_4_v;
    const double *bt_rem_u, *bt_rem_v, *IareaT, *IdxCu, *IdyCv;
    double *eta, *eta_pred, *ubt, *vbt, *ubt_prev, *vbt_prev, *uhbt, *vhbt;
    double *PFu, *PFv, *Cor_u, *Cor_v, *ubt_av, *vbt_av;
    const double *eta_in, *ubt_in, *vbt_in;
    double *u_av, *v_av, *eta_av;
};

__device__ inline double F4(const double *f, const Dims &d, int c, int i, int j)
{
    return f[4 * IX2(d, i, j) + (c - 1)];
}

// update_u at one point (:307-330)
__device__ inline void upd_u(const BT &p, int i, int j)
{
    const Dims &d = p.d;
    const long long c = IX2(d, i, j);
    const double Cor = (((F4(p.f_4_u, d, 4, i, j) * p.vbt[IX2(d, i + 1, j)]) + (F4(p.f_4_u, d, 1, i, j) * p.vbt[IX2(d, i, j - 1)])) +
                        ((F4(p.f_4_u, d, 3, i, j) * p.vbt[c]) + (F4(p.f_4_u, d, 2, i, j) * p.vbt[IX2(d, i + 1, j - 1)])));
    p.Cor_u[c] = Cor;
    p.ubt[c] = p.bt_rem_u[c] * (p.ubt[c] + p.dtbt * (Cor + p.PFu[c]));
}

// update_v at one point (:333-356)
__device__ inline void upd_v(const BT &p, int i, int j)
{
    const Dims &d = p.d;
    const long long c = IX2(d, i, j);
    const double Cor = -1.0 * (((F4(p.f_4_v, d, 1, i, j) * p.ubt[IX2(d, i - 1, j)]) + (F4(p.f_4_v, d, 4, i, j) * p.ubt[IX2(d, i, j + 1)])) +
                               ((F4(p.f_4_v, d, 2, i, j) * p.ubt[c]) + (F4(p.f_4_v, d, 3, i, j) * p.ubt[IX2(d, i - 1, j + 1)])));
    p.Cor_v[c] = Cor;
    p.vbt[c] = p.bt_rem_v[c] * (p.vbt[c] + p.dtbt * (Cor + p.PFv[c]));
}

__global__ void __launch_bounds__(256) k_btstep(BT p)
{
    cg::grid_group grid = cg::this_grid();
    const Dims &d = p.d;
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nth = (long long)gridDim.x * blockDim.x;
    const long long nall = d.s2;
    // points of the box is-1..ie+1 x js-1..je+1, enumerated row-major
    const int bw = (ie + 1) - (is - 1) + 1, bh = (je + 1) - (js - 1) + 1;
    const long long nbox = (long long)bw * bh;

    for (long long q = tid; q < nall; q += nth) {  // initialise from input (:199-207)
        p.eta[q] = p.eta_in[q];
        p.ubt[q] = p.ubt_in[q];
        p.vbt[q] = p.vbt_in[q];
        p.ubt_av[q] = 0.0;
        p.vbt_av[q] = 0.0;
    }
    grid.sync();

    const double trans_wt1 = 1.0 + p.bebt, trans_wt2 = -p.bebt;
    const double inv_nstep = 1.0 / (double)p.nstep;
    const int iu0 = p.extended ? is - 1 : is, iu1 = p.extended ? ie : ie - 1;
    const int jv0 = p.extended ? js - 1 : js, jv1 = p.extended ? je : je - 1;
    const int P = ie - is + 1, Q = je - js + 1;

    for (int n = 1; n <= p.nstep; ++n) {
        const bool v_first = ((n + p.first_direction) % 2) == 1;
        // P1 (:213-231)
        for (long long q = tid; q < nbox; q += nth) {
            const int i = (is - 1) + (int)(q % bw), j = (js - 1) + (int)(q / bw);
            const long long c = IX2(d, i, j);
            if (j >= js && j <= je) p.ubt_prev[c] = p.ubt[c];
            if (i >= is && i <= ie) p.vbt_prev[c] = p.vbt[c];
            if (i >= is && i <= ie && j >= js && j <= je)
                p.eta_pred[c] = p.eta[c] + (p.dtbt * p.IareaT[c]) *
                    (((p.Datu[IX2(d, i - 1, j)] * p.ubt[IX2(d, i - 1, j)]) - (p.Datu[c] * p.ubt[c])) +
                     ((p.Datv[IX2(d, i, j - 1)] * p.vbt[IX2(d, i, j - 1)]) - (p.Datv[c] * p.vbt[c])));
        }
        grid.sync();
        // P2 (:234-259): pressure forces, then the first update using them
        for (long long q = tid; q < nbox; q += nth) {
            const int i = (is - 1) + (int)(q % bw), j = (js - 1) + (int)(q / bw);
            const long long c = IX2(d, i, j);
            const bool in_u = i >= is && i <= ie - 1 && j >= js && j <= je;
            const bool in_v = i >= is && i <= ie && j >= js && j <= je - 1;
            if (in_u)
                p.PFu[c] = ((p.eta_pred[c] * p.gtot_E[c]) - (p.eta_pred[IX2(d, i + 1, j)] * p.gtot_W[IX2(d, i + 1, j)])) *
                           p.dgeo_de * p.IdxCu[c];
            if (in_v)
                p.PFv[c] = ((p.eta_pred[c] * p.gtot_N[c]) - (p.eta_pred[IX2(d, i, j + 1)] * p.gtot_S[IX2(d, i, j + 1)])) *
                           p.dgeo_de * p.IdyCv[c];
            if (v_first) {
                if (in_v) upd_v(p, i, j);
            } else {
                if (in_u) upd_u(p, i, j);
            }
        }
        grid.sync();
        // P3 (:252-272): second update, then this point's transports
        for (long long q = tid; q < nbox; q += nth) {
            const int i = (is - 1) + (int)(q % bw), j = (js - 1) + (int)(q / bw);
            const long long c = IX2(d, i, j);
            if (v_first) {
                if (i >= is && i <= ie - 1 && j >= js && j <= je) upd_u(p, i, j);
            } else {
                if (i >= is && i <= ie && j >= js && j <= je - 1) upd_v(p, i, j);
            }
            if (i >= iu0 && i <= iu1 && j >= js && j <= je)
                p.uhbt[c] = p.Datu[c] * (trans_wt1 * p.ubt[c] + trans_wt2 * p.ubt_prev[c]);
            if (i >= is && i <= ie && j >= jv0 && j <= jv1)
                p.vhbt[c] = p.Datv[c] * (trans_wt1 * p.vbt[c] + trans_wt2 * p.vbt_prev[c]);
        }
        grid.sync();
        // P4 (:274-291)
        for (long long q = tid; q < nbox; q += nth) {
            const int i = (is - 1) + (int)(q % bw), j = (js - 1) + (int)(q / bw);
            const long long c = IX2(d, i, j);
            if (i >= is && i <= ie && j >= js && j <= je)
                p.eta[c] = p.eta[c] - p.dtbt * p.IareaT[c] *
                    ((p.uhbt[c] - p.uhbt[IX2(d, i - 1, j)]) + (p.vhbt[c] - p.vhbt[IX2(d, i, j - 1)]));
            if (i >= is && i <= ie - 1 && j >= js && j <= je) p.ubt_av[c] = p.ubt_av[c] + p.ubt[c] * inv_nstep;
            if (i >= is && i <= ie && j >= js && j <= je - 1) p.vbt_av[c] = p.vbt_av[c] + p.vbt[c] * inv_nstep;
        }
        grid.sync();
        if ((p.px || p.py) && n < p.nstep) {
            // P5: halo refill; a halo cell copies its periodic image, which is
            // always a compute-domain cell (same result as x-then-y passes)
            for (long long q = tid; q < nall; q += nth) {
                const int i = d.isd + (int)(q % d.n1), j = d.jsd + (int)(q / d.n1);
                int si = i, sj = j;
                if (p.px) si = (i < is) ? i + P : (i > ie ? i - P : i);
                if (p.py) sj = (j < js) ? j + Q : (j > je ? j - Q : j);
                if (si == i && sj == j) continue;
                const long long s = IX2(d, si, sj);
                p.ubt[q] = p.ubt[s];
                p.vbt[q] = p.vbt[s];
                p.eta[q] = p.eta[s];
            }
            grid.sync();
        }
    }
    for (long long q = tid; q < nall; q += nth) {  // outputs (:296-302)
        p.u_av[q] = p.ubt_av[q];
        p.v_av[q] = p.vbt_av[q];
        p.eta_av[q] = p.eta[q];
    }
}

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_barotropic_init(mom6h_barotropic_cs **out, int isd, int ied, int jsd, int jed, int isc,
                          int iec, int jsc, int jec, int nstep, double dt, double bebt,
                          int first_direction, const double *Datu, const double *Datv,
                          const double *gtot_E, const double *gtot_W, const double *gtot_N,
                          const double *gtot_S, const double *f_4_u, const double *f_4_v,
                          const double *bt_rem_u, const double *bt_rem_v, const double *IareaT,
                          const double *IdxCu, const double *IdyCv, int periodic_x, int periodic_y)
{
    MOM6H_REQUIRE(out != nullptr, "mom6h_barotropic_init: NULL cs");
    MOM6H_REQUIRE(isd < isc && iec < ied && jsd < jsc && jec < jed && isc < iec && jsc < jec,
                  "mom6h_barotropic_init: needs one halo cell on every side");
    MOM6H_REQUIRE(nstep >= 1 && dt > 0.0, "mom6h_barotropic_init: nstep >= 1 and dt > 0 required");
    MOM6H_REQUIRE(!periodic_x || (iec - isc + 1) >= (isc - isd) && (iec - isc + 1) >= (ied - iec),
                  "mom6h_barotropic_init: compute domain narrower than the x halo");
    MOM6H_REQUIRE(!periodic_y || (jec - jsc + 1) >= (jsc - jsd) && (jec - jsc + 1) >= (jed - jec),
                  "mom6h_barotropic_init: compute domain narrower than the y halo");
    MOM6H_REQUIRE(Datu && Datv && gtot_E && gtot_W && gtot_N && gtot_S && f_4_u && f_4_v && bt_rem_u &&
                      bt_rem_v && IareaT && IdxCu && IdyCv,
                  "mom6h_barotropic_init: NULL array");
    int coop = 0, dev = 0;
    MOM6H_CHECK(hipGetDevice(&dev));
    MOM6H_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    MOM6H_REQUIRE(coop, "mom6h_barotropic_init: device lacks cooperative launch");
    auto *cs = new mom6h_barotropic_cs();
    cs->d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, 1);
    cs->halo = Halo{periodic_x, periodic_y};
    cs->nstep = nstep;
    cs->first_direction = first_direction;
    cs->dtbt = dt / (double)nstep;   // barotropic_init (:72-75)
    cs->bebt = bebt;
    cs->dgeo_de = 1.0;
    const size_t n2 = (size_t)cs->d.s2;
    int rc = 0;
    const double *src[] = {Datu, Datv, gtot_E, gtot_W, gtot_N, gtot_S, bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv};
    double **dst[] = {&cs->Datu, &cs->Datv, &cs->gtot_E, &cs->gtot_W, &cs->gtot_N, &cs->gtot_S,
                      &cs->bt_rem_u, &cs->bt_rem_v, &cs->IareaT, &cs->IdxCu, &cs->IdyCv};
    for (size_t q = 0; q < sizeof(src) / sizeof(src[0]); ++q) rc = rc ? rc : dev_upload(dst[q], src[q], n2);
    rc = rc ? rc : dev_upload(&cs->f_4_u, f_4_u, 4 * n2);
    rc = rc ? rc : dev_upload(&cs->f_4_v, f_4_v, 4 * n2);
    for (double **pp : {&cs->eta, &cs->eta_pred, &cs->ubt, &cs->vbt, &cs->ubt_prev, &cs->vbt_prev,
                        &cs->uhbt, &cs->vhbt, &cs->PFu, &cs->PFv, &cs->Cor_u, &cs->Cor_v,
                        &cs->ubt_av, &cs->vbt_av})
        rc = rc ? rc : dev_alloc(pp, n2, true);
    if (rc) {
        mom6h_barotropic_end(cs);
        return rc;
    }
    // persistent grid: as many resident 256-thread blocks as the device holds,
    // capped by the work of one phase
    int per_cu = 0, ncu = 0;
    MOM6H_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_btstep, 256, 0));
    MOM6H_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long long box = (long long)(iec - isc + 3) * (jec - jsc + 3);
    long long want = (box + 255) / 256;
    long long cap = (long long)per_cu * ncu;
    if (cap > (long long)ncu) cap = ncu;   // one block per CU leaves room for the 3D kernels
    cs->grid_blocks = (int)(want < cap ? want : cap);
    if (cs->grid_blocks < 1) cs->grid_blocks = 1;
    *out = cs;
    return MOM6H_OK;
}

int mom6h_btstep(mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in,
                 const double *vbt_in, double *u_av, double *v_av, double *eta_av, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_btstep: NULL cs");
    MOM6H_REQUIRE(eta_in && ubt_in && vbt_in && u_av && v_av && eta_av, "mom6h_btstep: NULL array");
    BT p;
    p.d = cs->d;
    p.nstep = cs->nstep; p.first_direction = cs->first_direction;
    p.px = cs->halo.px; p.py = cs->halo.py;
    p.extended = (p.px || p.py) ? 1 : 0;
    p.dtbt = cs->dtbt; p.bebt = cs->bebt; p.dgeo_de = cs->dgeo_de;
    p.Datu = cs->Datu; p.Datv = cs->Datv; p.gtot_E = cs->gtot_E; p.gtot_W = cs->gtot_W;
    p.gtot_N = cs->gtot_N; p.gtot_S = cs->gtot_S; p.f_4_u = cs->f_4_u; p.f_4_v = cs->f_4_v;
    p.bt_rem_u = cs->bt_rem_u; p.bt_rem_v = cs->bt_rem_v; p.IareaT = cs->IareaT;
    p.IdxCu = cs->IdxCu; p.IdyCv = cs->IdyCv;
    p.eta = cs->eta; p.eta_pred = cs->eta_pred; p.ubt = cs->ubt; p.vbt = cs->vbt;
    p.ubt_prev = cs->ubt_prev; p.vbt_prev = cs->vbt_prev; p.uhbt = cs->uhbt; p.vhbt = cs->vhbt;
    p.PFu = cs->PFu; p.PFv = cs->PFv; p.Cor_u = cs->Cor_u; p.Cor_v = cs->Cor_v;
    p.ubt_av = cs->ubt_av; p.vbt_av = cs->vbt_av;
    p.eta_in = eta_in; p.ubt_in = ubt_in; p.vbt_in = vbt_in;
    p.u_av = u_av; p.v_av = v_av; p.eta_av = eta_av;
    void *args[] = {&p};
    MOM6H_CHECK(hipLaunchCooperativeKernel((const void *)k_btstep, dim3(cs->grid_blocks), dim3(256),
                                           args, 0, (hipStream_t)stream));
    return MOM6H_OK;
}

int mom6h_barotropic_end(mom6h_barotropic_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->Datu, cs->Datv, cs->gtot_E, cs->gtot_W, cs->gtot_N, cs->gtot_S, cs->f_4_u,
                      cs->f_4_v, cs->bt_rem_u, cs->bt_rem_v, cs->IareaT, cs->IdxCu, cs->IdyCv,
                      cs->eta, cs->eta_pred, cs->ubt, cs->vbt, cs->ubt_prev, cs->vbt_prev, cs->uhbt,
                      cs->vhbt, cs->PFu, cs->PFv, cs->Cor_u, cs->Cor_v, cs->ubt_av, cs->vbt_av})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

}  // extern "C"
