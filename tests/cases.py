"""Configurations shared by the parity tests (small enough for the oracle to
finish in well under a second each)."""
from hackathon_mom6_miniapp_amd.state import BC_PERIODIC, BC_STATIC, BC_WALL, Config

# name -> (Config, init, niter)
CASES = {
    "static_ref_16": (Config(ni=16, nj=16, nk=5, bt_nsteps=10), "reference", 2),
    "static_seeded_ragged": (Config(ni=22, nj=15, nk=6, bt_nsteps=12), "seeded", 3),
    "doubly_periodic_fplane": (Config(ni=24, nj=24, nk=8, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC,
                                      coriolis="fplane", f0=1e-4), "seeded", 3),
    "zonal_periodic_walls_beta": (Config(ni=28, nj=20, nk=7, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                         coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3), "seeded", 3),
    "biharmonic_walls": (Config(ni=24, nj=24, nk=6, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, biharmonic=True, Ah=1e11),
                         "seeded", 2),
    "cw84_limiter": (Config(ni=20, nj=20, nk=6, bt_nsteps=15, monotonic=True), "seeded", 2),
    # no-slip walls with the biharmonic term: the tiled hor_visc kernel's
    # no-slip and Del2 paths (the closure cases run the general kernel)
    "noslip_biharmonic_walls": (Config(ni=70, nj=24, nk=5, bt_nsteps=12, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                       coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, biharmonic=True,
                                       Ah=1e11, no_slip=True), "seeded", 2),
    "arakawa_hsu90": (Config(ni=20, nj=20, nk=6, bt_nsteps=15, coriolis_scheme=2), "reference", 2),
    "arakawa_lamb81": (Config(ni=20, nj=20, nk=6, bt_nsteps=15, coriolis_scheme=3), "seeded", 2),
    "deep_column_64": (Config(ni=64, nj=48, nk=30, bt_nsteps=30), "reference", 2),
    # Rayleigh drag at every level (the 3D vertvisc_type fields), pinned by its golden fixture
    "ray3d_walls_beta": (Config(ni=24, nj=20, nk=7, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, ray_bottom=1e-3,
                                ray_profile="linear"), "seeded", 3),
    "ray3d_seams_deep": (Config(ni=140, nj=90, nk=40, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, ray_bottom=1e-3,
                                ray_profile="linear"), "seeded", 2),
}

# fields compared after the run in the module-by-module mode: every array the
# reference driver holds
COMPARE = ("u", "v", "h", "htmp", "up", "vp", "uh", "vh", "CAu", "CAv", "diffu", "diffv",
           "eta", "ubt_av", "vbt_av", "eta_av", "u_cor", "du_cor",
           "FA_u_EE", "FA_u_E0", "FA_u_W0", "FA_u_WW", "uBT_WW", "uBT_EE", "BT_h_u",
           "vv_taux_bot", "vv_tauy_bot")

# The fused path keeps the transports, tendencies and Coriolis terms on chip:
# everything else is compared whole, uh on the continuity ranges (the only
# cells the final continuity call writes).
COMPARE_FUSED = ("u", "v", "h", "htmp", "up", "vp", "eta", "ubt_av", "vbt_av", "eta_av", "u_cor", "du_cor",
                 "FA_u_EE", "FA_u_E0", "FA_u_W0", "FA_u_WW", "uBT_WW", "uBT_EE", "BT_h_u",
                 "vv_taux_bot", "vv_tauy_bot")

# continuity with the meridional pass after the zonal one (no reference
# counterpart; the oracle's meridional pass is its zonal pass on the
# transposed grid)
MERID_CASES = {
    "merid_zonal_periodic_walls_beta": (Config(ni=28, nj=20, nk=7, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                               coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3,
                                               meridional=True), "seeded", 3),
    "merid_doubly_periodic_fplane": (Config(ni=24, nj=26, nk=8, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC,
                                            coriolis="fplane", f0=1e-4, meridional=True), "seeded", 2),
    "merid_static_ref": (Config(ni=20, nj=20, nk=5, bt_nsteps=10, meridional=True), "reference", 2),
}
MERID_COMPARE = ("v_cor", "dv_cor", "FA_v_NN", "FA_v_N0", "FA_v_S0", "FA_v_SS", "vBT_SS", "vBT_NN", "BT_h_v")

# hor_visc closures (Smagorinsky, Leith, thickness-aware bounds, no-slip):
# the general hor_visc kernel, module by module (the fused stencil has no
# closures); the same configurations as the golden fixtures that pin the oracle
CLOSURE_CASES = {
    "smag_bounds_noslip_walls": (Config(ni=24, nj=20, nk=6, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                        coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, biharmonic=True, Ah=1e11,
                                        smagorinsky=True, no_slip=True, better_bound=True), "seeded", 3),
    "smag_laplacian_bound_periodic": (Config(ni=22, nj=22, nk=5, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC,
                                             coriolis="fplane", f0=1e-4, Kh=20.0, smagorinsky=True,
                                             better_bound=True), "seeded", 3),
    "leith_biharmonic_walls": (Config(ni=24, nj=20, nk=6, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                      coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=10.0, biharmonic=True, Ah=1e9,
                                      leith=True), "seeded", 3),
    "leith_smag_noslip_static": (Config(ni=20, nj=20, nk=6, bt_nsteps=15, Kh=10.0, smagorinsky=True, leith=True,
                                        no_slip=True), "reference", 2),
    # across tile seams, all closures at once
    "closures_seams_200x140": (Config(ni=200, nj=140, nk=8, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                      coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, biharmonic=True, Ah=1e11,
                                      smagorinsky=True, leith=True, no_slip=True, better_bound=True), "seeded", 2),
    # no-slip alone (the constant-viscosity kernels)
    "noslip_walls_biharmonic": (Config(ni=24, nj=24, nk=6, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                       coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, biharmonic=True, Ah=1e11,
                                       no_slip=True), "seeded", 2),
}

# the coupled form (continuity's barotropic-transport targets from btstep):
# the golden configurations (pinned by the reference's continuity_PPM with
# uhbt present) and seams, with a Gaussian sea-surface bump so that the
# barotropic solution is not zero.  (Config, init, niter, bump)
COUPLED_CASES = {
    "coupled_walls_beta": (Config(ni=24, nj=20, nk=7, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                  coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, coupled=True),
                           "seeded", 3, (0.5, 40.0e3)),
    "coupled_static_reference": (Config(ni=20, nj=20, nk=6, bt_nsteps=15, coupled=True), "reference", 3, None),
    "coupled_periodic_fplane_merid": (Config(ni=24, nj=24, nk=6, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC,
                                             coriolis="fplane", f0=1e-4, coupled=True, meridional=True),
                                      "seeded", 3, (0.5, 40.0e3)),
    "coupled_seams_200x140": (Config(ni=200, nj=140, nk=12, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                     coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, coupled=True),
                              "seeded", 2, (0.5, 100.0e3)),
    "coupled_seams_deep_merid": (Config(ni=140, nj=100, nk=40, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                        coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, coupled=True,
                                        meridional=True), "seeded", 2, (0.5, 100.0e3)),
}
