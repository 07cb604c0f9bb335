"""Small setups that together exercise every plugin class the engine lowers.

Each builder returns a configs.Setup whose trajectories are short (a few thousand steps), so the
oracle finishes a few hundred of them in seconds.  The GPU parity test runs every setup through the
FP64 kernel and the oracle with the same counter streams.
"""
import numpy as np

from spectrum_amd import configs as cf
from spectrum_amd.data_container import DataContainer as DC


def dist(n_bins, lo, hi, log=(0, 0, 0), outside=(0, 0, 0)):
    return cf.distribution_base(DC(), n_bins, lo, hi, log, outside)


def shock_flowmom_reflect():
    """Forward Parker at a discontinuous shock; kappa ~ |U|^a p^b (position dependent through U);
    log-uniform initial momenta; a reflecting plane upstream; time pass + expiry; (x, p) map."""
    s = cf.Setup("shock_flowmom_reflect", cf.TRAJ_PARKER, cf.FORWARD)
    u1, B = 0.013343, 16.0
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (u1, 0.0, 0.0), (B, 0.0, 0.0), 1.0)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.0).insert_double(3.0)
    s.add(cf.BACKGROUND, "BackgroundShock", dc)
    s.add(cf.DIFFUSION, "DiffusionFlowMomentumPowerLaw",
          DC().insert_double(3.0).insert_double(u1).insert_double(0.5).insert_double(0.05).insert_double(0.3).insert_double(0.1))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((-3.0, 1.0, -2.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(0.03).insert_double(0.3).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (-30.0, 0.0, 0.0)), ("vec", (1.0, 0.0, 0.0))))
    s.add(cf.BOUNDARY, "BoundaryTimePass", cf.boundary(-1, [0], ("double", 150.0)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 400.0)))
    d = dist((12, 0, 10), (-36.0, 0.0, 0.02), (12.0, 0.0, 1.0), (0, 0, 1), (1, 0, 1))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(0).insert_int(0)
    s.add(cf.DISTRIBUTION, "DistributionPositionMomentumUniform", d)
    return s


def solarwind_kinrad_reflect_recurrent():
    """Backward Parker in the Parker spiral; kappa ~ T^a r^b (position dependent); a reflecting inner
    sphere; a recurrent time boundary binning every 12 h; exit-time and <dx> distributions."""
    s = cf.Setup("solarwind_kinrad_reflect_recurrent", cf.TRAJ_PARKER, cf.BACKWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    kd = DC().insert_double(1.5e22 / cf.unit_diffusion).insert_double(cf.GeV).insert_double(cf.one_au)
    kd.insert_double(1.0).insert_double(1.0).insert_double(0.05).insert_int(0).insert_double(1.0).insert_double(0.0).insert_double(1.0)
    s.add(cf.DIFFUSION, "DiffusionKineticEnergyRadialDistancePowerLaw", kd)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.2 * cf.one_au, 0.3 * cf.one_au, -0.1 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(0.3 * cf.GeV)))
    s.add(cf.BOUNDARY, "BoundarySphereReflect", cf.boundary(-1, [-1, -1], ("vec", (0.0, 0.0, 0.0)), ("double", 1.1 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeRecurrent", cf.boundary(-1, [0, 0], ("double", -0.5 * cf.one_day)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0, 0], ("double", -2.25 * cf.one_day)))
    d = dist((10, 0, 0), (-2.5 * cf.one_day, 0, 0), (0.0, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    s.add(cf.DISTRIBUTION, "DistributionPositionCumulativeOrder1", dist((8, 0, 0), (-2.5 * cf.one_day, 0, 0), (0.0, 0, 0)))
    return s


def parker_offset_spheres():
    """Backward Parker in the Parker spiral (the headline's diffusion) with three spheres: an absorbing one of
    3 au and a reflecting one of 0.6 au around the origin (consecutive slots: the second reuses the first's
    distance), and an absorbing one of 0.15 au around (0.75, 0.1, 0) au (not concentric with the slot
    before it); a one-day limit, exit-time and spectrum distributions."""
    s = cf.Setup("parker_offset_spheres", cf.TRAJ_PARKER, cf.BACKWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    lam0 = 3.0 * (1.5e22 / cf.unit_diffusion) / cf.c_code
    dif = DC().insert_double(lam0).insert_double(cf.GV).insert_double(5.0e-5 / cf.unit_magnetic)
    dif.insert_double(0.5).insert_double(0.0).insert_double(0.02)
    s.add(cf.DIFFUSION, "DiffusionRigidityMagneticFieldPowerLaw", dif)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((cf.one_au, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(0.05 * cf.GeV)).insert_double(cf.mom_from_ekin(2.0 * cf.GeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0, 0], ("vec", (0.0, 0.0, 0.0)), ("double", 3.0 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundarySphereReflect", cf.boundary(-1, [-1, -1], ("vec", (0.0, 0.0, 0.0)), ("double", 0.6 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0, 0], ("vec", (0.75 * cf.one_au, 0.1 * cf.one_au, 0.0)), ("double", 0.15 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0, 0], ("double", -1.0 * cf.one_day)))
    d = dist((10, 0, 0), (-1.0 * cf.one_day, 0, 0), (0.0, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    s.add(cf.DISTRIBUTION, "DistributionPositionCumulativeOrder1", dist((8, 0, 0), (-1.0 * cf.one_day, 0, 0), (0.0, 0, 0)))
    return s


def smoothshock_perp_momentum():
    """Forward ParkerSource at a smooth shock with B tangential; perpendicular-only diffusion;
    momentum pass and injection boundaries; <dx dx> and (p) distributions."""
    s = cf.Setup("smoothshock_perp_momentum", cf.TRAJ_PARKER_SOURCE, cf.FORWARD)
    u1, w = 0.013343, 2.5
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (u1, 0.0, 0.0), (0.0, 16.0, 0.0), 10.0 * w)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.0).insert_double(4.0).insert_double(w).insert_double(0.1)
    s.add(cf.BACKGROUND, "BackgroundSmoothShock", dc)
    s.add(cf.DIFFUSION, "DiffusionPerpConstant", DC().insert_double(3.3356))
    p0 = cf.mom_from_ekin(10.0 * cf.keV)
    s.add(cf.SOURCE, "SourceMomentumInjection", DC().insert_double(1.3 * p0).insert_double(2.0))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(p0))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0, 0], ("double", 20000.0)))
    s.add(cf.BOUNDARY, "BoundaryMomentumPass", cf.boundary(-1, [0, -1], ("double", 1.2 * p0)))
    s.add(cf.BOUNDARY, "BoundaryMomentumInject", cf.boundary(1, [0, 0], ("double", 1.6 * p0)))
    s.add(cf.DISTRIBUTION, "DistributionPositionCumulativeOrder2", dist((6, 0, 0), (0.0, 0, 0), (21000.0, 0, 0)))
    d = dist((16, 1, 1), (p0, -1.0, -1.0), (2.0 * p0, 1.0, 1.0), (1, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(0)
    s.add(cf.DISTRIBUTION, "DistributionMomentumUniform", d)
    return s


def uniform_para_plane_pass():
    """Forward Parker in a uniform field with parallel-only diffusion and a pass-through plane."""
    s = cf.Setup("uniform_para_plane_pass", cf.TRAJ_PARKER, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.002, 0.001, 0.0), (3.0, 4.0, 12.0), 5.0))
    s.add(cf.DIFFUSION, "DiffusionParaConstant", DC().insert_double(2.5))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumFixed", DC().insert_vec((cf.mom_from_ekin(cf.MeV), 0.0, 0.0)))
    s.add(cf.BOUNDARY, "BoundaryPlanePass", cf.boundary(-1, [0], ("vec", (0.0, 0.0, 20.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 600.0)))
    d = dist((12, 0, 0), (0.0, 0, 0), (660.0, 0, 0))
    d.insert_double(1.0).insert_double(0.5).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def modulation_spectrum_type(t):
    s = cf.parker_modulation(n_energies=8, t_limit_years=2.0 / 365.0, time_action=0, spectrum_type=t, name=f"modulation_type{t}")
    return s


def gc_rigidity_reflect():
    """Guiding centre in the Parker spiral with a |B|-dependent rigidity law (perpendicular drift from
    grad kappa), no pitch-angle scattering model term, a reflecting sphere and a (|p|, mu) map."""
    s = cf.Setup("gc_rigidity_reflect", cf.TRAJ_GUIDING_DIFF_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    lam0 = 3.0 * (1.5e22 / cf.unit_diffusion) / cf.c_code
    dif = DC().insert_double(lam0).insert_double(cf.GV).insert_double(5.0e-5 / cf.unit_magnetic)
    dif.insert_double(0.5).insert_double(-1.0).insert_double(0.02)
    s.add(cf.DIFFUSION, "DiffusionRigidityMagneticFieldPowerLaw", dif)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((cf.one_au, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(50.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundarySphereReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, 0.0)), ("double", 0.9995 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.2 * cf.one_day)))
    d = dist((8, 8, 1), (4.0, -1.0, -1.0), (4.3, 1.0, 1.0), (0, 0, 0), (1, 1, 1))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionMomentumUniform", d)
    return s


def gc_isotropic_only():
    """Guiding centre with pitch-angle scattering only (no perpendicular diffusion)."""
    s = cf.Setup("gc_isotropic_only", cf.TRAJ_GUIDING_DIFF_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.DIFFUSION, "DiffusionIsotropicConstant", DC().insert_double(5.0e-3))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((2.0 * cf.one_au, 0.5 * cf.one_au, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumRing", DC().insert_double(cf.mom_from_ekin(20.0 * cf.MeV)).insert_double(0.4))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.1 * cf.one_day)))
    d = dist((16, 16, 0), (1.5 * cf.one_au, -1.0, 0.0), (3.0 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def guiding_spiral():
    """TrajectoryGuiding (no diffusion): gradient and curvature drifts of a 20 MeV ring in the Parker
    spiral, no diffusion model, a time limit and a (|p|, mu) map."""
    s = cf.Setup("guiding_spiral", cf.TRAJ_GUIDING, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.5 * cf.one_au, 0.3 * cf.one_au, 0.2 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(20.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.05 * cf.one_day)))
    d = dist((8, 8, 1), (3.0, -1.0, -1.0), (3.5, 1.0, 1.0), (0, 0, 0), (1, 1, 1))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionMomentumUniform", d)
    return s


def guiding_scatt_uniform():
    """TrajectoryGuidingScatt: isotropic pitch-angle scattering in a uniform field along z, no flow."""
    s = cf.Setup("guiding_scatt_uniform", cf.TRAJ_GUIDING_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 3.0), 50.0))
    s.add(cf.DIFFUSION, "DiffusionIsotropicConstant", DC().insert_double(2.0e-3))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.0, 2.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumRing", DC().insert_double(cf.mom_from_ekin(10.0 * cf.MeV)).insert_double(0.6))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 300.0)))
    d = dist((16, 16, 0), (-500.0, -1.0, 0.0), (500.0, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def guiding_diff_rigidity():
    """TrajectoryGuidingDiff: perpendicular diffusion with a |B|-dependent rigidity law (grad kappa drift)
    in the Parker spiral, a reflecting sphere, no pitch-angle scattering."""
    s = gc_rigidity_reflect()
    s.name, s.traj_type = "guiding_diff_rigidity", cf.TRAJ_GUIDING_DIFF
    return s


def parker_strauss_heliopause():
    """Parker (backward) on the SolarWind background across the 117 au heliopause with the Strauss et al. 2013
    model selecting inner / outer parameters by region[0], a sphere at 100 au and a spectrum."""
    s = cf.Setup("parker_strauss_heliopause", cf.TRAJ_PARKER, cf.BACKWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    p0 = cf.mom_from_ekin(1.0 * cf.GeV)
    dc = DC().insert_int(0).insert_double(0.5 * cf.one_au).insert_double(2.0 * cf.one_au).insert_double(p0 * cf.c_code)
    dc.insert_double(5.0e-6 / cf.unit_magnetic).insert_double(0.02).insert_double(0.2)
    s.add(cf.DIFFUSION, "DiffusionStraussEtAl2013", dc)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.6 * 116.8 * cf.one_au, 0.64 * 116.8 * cf.one_au, 0.48 * 116.8 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(0.3 * p0).insert_double(3.0 * p0).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 100.0 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", -5.0 * cf.one_day)))
    d = dist((8, 1, 1), (0.1 * p0, -1.0, -1.0), (10.0 * p0, 1.0, 1.0), (1, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(0)
    s.add(cf.DISTRIBUTION, "DistributionMomentumUniform", d)
    return s


def parker_potgieter_uniform():
    """Parker (forward) in a uniform flow and field with the Potgieter et al. 2015 model (no region) and a
    reflecting plane."""
    s = cf.Setup("parker_potgieter_uniform", cf.TRAJ_PARKER, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.002, 0.0, 0.001), (2.0, 1.0, 2.0), 0.5))
    p0 = cf.mom_from_ekin(50.0 * cf.MeV)
    dc = DC().insert_int(-1).insert_double(2.0).insert_double(20.0).insert_double(p0 * cf.c_code)
    dc.insert_double(4.5).insert_double(0.05).insert_double(0.5)
    s.add(cf.DIFFUSION, "DiffusionPotgieterEtAl2015", dc)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(0.5 * p0).insert_double(2.0 * p0).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (-3.0, 0.0, 0.0)), ("vec", (1.0, 0.0, 0.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 200.0)))
    s.add(cf.DISTRIBUTION, "DistributionPositionCumulativeOrder1", dist((10, 0, 0), (0.0, 0, 0), (201.0, 0, 0)))
    return s


def gc_qlt_spiral():
    """Guiding centre (GuidingDiffScatt) in the Parker spiral with QLT pitch-angle scattering, whose mu
    derivative is DiffusionBase's forward difference."""
    s = cf.Setup("gc_qlt_spiral", cf.TRAJ_GUIDING_DIFF_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.DIFFUSION, "DiffusionQLTConstant", DC().insert_double(0.05).insert_double(0.01 * cf.one_au).insert_double(5.0 / 3.0))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.2 * cf.one_au, 0.4 * cf.one_au, 0.1 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(20.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.01 * cf.one_day)))
    d = dist((16, 16, 0), (1.0 * cf.one_au, -1.0, 0.0), (1.5 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def fieldline_spiral():
    """TrajectoryFieldline (forward) along the Parker spiral from 1 au off the ecliptic: no diffusion model, a
    reflecting plane the line crosses (the momentum reverses and the centre runs back inward), absorbing
    spheres at 0.5 and 3 au and a time distribution."""
    s = cf.Setup("fieldline_spiral", cf.TRAJ_FIELDLINE, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.8 * cf.one_au, 0.5 * cf.one_au, 0.2 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_double(cf.mom_from_ekin(100.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, 0.35 * cf.one_au)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 3.0 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 0.5 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 2.0 * cf.one_day)))
    d = dist((12, 0, 0), (0.0, 0, 0), (4.0 * cf.one_day, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def fieldline_backward_shock():
    """TrajectoryFieldline (backward) through a smooth shock with an oblique field."""
    s = cf.Setup("fieldline_backward_shock", cf.TRAJ_FIELDLINE, cf.BACKWARD)
    u1, w = 0.013343, 2.5
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (u1, 0.0, 0.0), (8.0, 6.0, 0.0), 2.0 * w)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.0).insert_double(4.0).insert_double(w).insert_double(0.1)
    s.add(cf.BACKGROUND, "BackgroundSmoothShock", dc)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((20.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", -600.0)))
    d = dist((10, 0, 0), (-600.0, 0, 0), (0.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def lorentz_uniform():
    """TrajectoryLorentz (forward) in a uniform oblique field carried by a flow (E = -u x B / c): isotropic
    momenta on a thick shell, a reflecting plane below the start, a pass plane above it and a time limit.
    The time step is the gyro period / 100 for the slow particles and 0.1 dmax / v for the fast ones."""
    s = cf.Setup("lorentz_uniform", cf.TRAJ_LORENTZ, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.002, 0.001, 0.0), (3.0, 4.0, 12.0), 1.0))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_double(cf.mom_from_ekin(100.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, -3.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryPlanePass", cf.boundary(-1, [0], ("vec", (0.0, 0.0, 6.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 40.0)))
    d = dist((10, 0, 0), (0.0, 0, 0), (44.0, 0, 0))
    d.insert_double(1.0).insert_double(0.5).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def lorentz_spiral_backward():
    """TrajectoryLorentz (backward) in the solar wind near 1 au: protons on a shell, the motional field of
    the flow at the start point, an absorbing sphere just inside the start and a time limit."""
    s = cf.Setup("lorentz_spiral_backward", cf.TRAJ_LORENTZ, cf.BACKWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.8 * cf.one_au, 0.5 * cf.one_au, 0.2 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(10.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 0.9635 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", -30.0)))
    d = dist((10, 0, 0), (-30.0, 0, 0), (0.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def focused_spiral():
    """TrajectoryFocused (forward) along the Parker spiral from 1 au off the ecliptic: adiabatic focusing,
    adiabatic cooling in the expanding wind and the gradient/curvature drifts; a reflecting plane, absorbing
    spheres at 0.5 and 3 au, a time limit, a momentum and a time distribution."""
    s = cf.Setup("focused_spiral", cf.TRAJ_FOCUSED, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.8 * cf.one_au, 0.5 * cf.one_au, 0.2 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_double(cf.mom_from_ekin(100.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1, -1], ("vec", (0.0, 0.0, 0.35 * cf.one_au)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0, 0], ("vec", (0.0, 0.0, 0.0)), ("double", 3.0 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0, 0], ("vec", (0.0, 0.0, 0.0)), ("double", 0.5 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0, 0], ("double", 2.0 * cf.one_day)))
    d = dist((12, 0, 0), (0.0, 0, 0), (4.0 * cf.one_day, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    d = dist((10, 8, 1), (cf.mom_from_ekin(0.5 * cf.MeV), -1.0, -1.0), (cf.mom_from_ekin(120.0 * cf.MeV), 1.0, 1.0), (1, 0, 0), (1, 1, 1))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionMomentumUniform", d)
    return s


def focused_uniform_backward():
    """TrajectoryFocused (backward) in a uniform oblique field carried by a flow: a ring of pitch cosine 0.6,
    straight drift at U + v mu b through a pass plane, against a reflecting plane, to a time limit."""
    s = cf.Setup("focused_uniform_backward", cf.TRAJ_FOCUSED, cf.BACKWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.02, 0.01, 0.0), (3.0, 4.0, 12.0), 2.0))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumRing", DC().insert_double(cf.mom_from_ekin(10.0 * cf.MeV)).insert_double(0.6))
    s.add(cf.BOUNDARY, "BoundaryPlanePass", cf.boundary(-1, [0], ("vec", (0.0, 0.0, -20.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, -60.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", -200.0)))
    d = dist((10, 0, 0), (-220.0, 0, 0), (0.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def wnlt(ramp=False):
    """DiffusionWNLTConstant (A2A, l_max, ps_index, A2T, A2L) or DiffusionWNLTRampVLISM (+ l_max_HP, z_nose,
    z_sheath) with a Kolmogorov spectrum."""
    dc = DC().insert_double(0.05).insert_double(0.01 * cf.one_au).insert_double(5.0 / 3.0).insert_double(0.1).insert_double(0.2)
    if ramp:
        dc.insert_double(0.002 * cf.one_au).insert_double(0.5 * cf.one_au).insert_double(1.5 * cf.one_au)
    return ("DiffusionWNLTRampVLISM" if ramp else "DiffusionWNLTConstant"), dc


def gc_wnlt_scatt_spiral():
    """GuidingScatt in the Parker spiral with WNLT pitch-angle scattering (QLT plus the nonlinear 2F1 term;
    the mu derivative is DiffusionBase's forward difference)."""
    s = cf.Setup("gc_wnlt_scatt_spiral", cf.TRAJ_GUIDING_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.DIFFUSION, *wnlt())
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.2 * cf.one_au, 0.4 * cf.one_au, 0.1 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(20.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.0001 * cf.one_day)))
    d = dist((16, 16, 0), (1.0 * cf.one_au, -1.0, 0.0), (1.5 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def gc_wnlt_ramp_diffscatt():
    """GuidingDiffScatt in the Parker spiral with the WNLT ramp model: kappa_perp from WNLT and its
    perpendicular drift by DiffusionBase's directional derivative over the solar wind's derivative steps."""
    s = cf.Setup("gc_wnlt_ramp_diffscatt", cf.TRAJ_GUIDING_DIFF_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.DIFFUSION, *wnlt(ramp=True))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.2 * cf.one_au, 0.4 * cf.one_au, 0.1 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(5.0 * cf.MeV)).insert_double(cf.mom_from_ekin(50.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.00015 * cf.one_day)))
    d = dist((16, 16, 0), (1.0 * cf.one_au, -1.0, 0.0), (1.5 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def gc_strauss_diff():
    """GuidingDiff in the Parker spiral with the Strauss et al. 2013 model (inner parameters by region[0]):
    kappa_perp ~ 1/|B| and its perpendicular drift by DiffusionBase's directional derivative."""
    s = cf.Setup("gc_strauss_diff", cf.TRAJ_GUIDING_DIFF, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    p0 = cf.mom_from_ekin(1.0 * cf.GeV)
    dc = DC().insert_int(0).insert_double(0.002 * cf.one_au).insert_double(2.0 * cf.one_au).insert_double(p0 * cf.c_code)
    dc.insert_double(5.0e-6 / cf.unit_magnetic).insert_double(0.05).insert_double(0.2)
    s.add(cf.DIFFUSION, "DiffusionStraussEtAl2013", dc)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.2 * cf.one_au, 0.4 * cf.one_au, 0.1 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(5.0 * cf.MeV)).insert_double(cf.mom_from_ekin(500.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 1.4 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 0.3 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.1 * cf.one_day)))
    d = dist((16, 16, 0), (1.0 * cf.one_au, -1.0, 0.0), (1.5 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def focused_smoothshock():
    """TrajectoryFocused (forward) through a moving smooth shock with an oblique field: analytic grad U and
    grad B across the transition and dU/dt of the moving shock; the dmax of the shock's hidden coordinate."""
    s = cf.Setup("focused_smoothshock", cf.TRAJ_FOCUSED, cf.FORWARD)
    u1, w = 0.013343, 2.5
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (u1, 0.0, 0.0), (8.0, 6.0, 0.0), 2.0 * w)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.002).insert_double(4.0).insert_double(w).insert_double(0.1)
    s.add(cf.BACKGROUND, "BackgroundSmoothShock", dc)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((-4.0, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(0.1 * cf.MeV)).insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneAbsorb", cf.boundary(1, [0], ("vec", (-40.0, 0.0, 0.0)), ("vec", (1.0, 0.0, 0.0))))
    s.add(cf.BOUNDARY, "BoundaryPlaneAbsorb", cf.boundary(1, [0], ("vec", (40.0, 0.0, 0.0)), ("vec", (1.0, 0.0, 0.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 300.0)))
    d = dist((10, 0, 0), (0.0, 0, 0), (330.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def uniform_sphere_start():
    """Forward Parker in a uniform flow and field started on a sphere (InitialSpaceSphere, two uniforms
    before the momentum's), an absorbing sphere further out and a time limit; (r, p) map."""
    s = cf.Setup("uniform_sphere_start", cf.TRAJ_PARKER, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundUniform", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.01, -0.005, 0.002), (2.0, 1.0, 3.0), 0.4))
    s.add(cf.DIFFUSION, "DiffusionFullConstant", DC().insert_double(0.5).insert_double(2.0))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceSphere", DC().insert_vec((1.0, 2.0, 3.0)).insert_double(5.0))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(0.05).insert_double(0.5).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (1.0, 2.0, 3.0)), ("double", 12.0)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [1], ("double", 60.0)))
    d = dist((10, 0, 8), (0.0, 0.0, 0.05), (15.0, 0.0, 0.5), (0, 0, 1), (1, 0, 1))
    d.insert_double(1.0).insert_double(0.5).insert_int(1).insert_int(3).insert_int(0)
    s.add(cf.DISTRIBUTION, "DistributionPositionMomentumUniform", d)
    return s


def gc_sector_start():
    """Guiding centre with perpendicular diffusion and pitch-angle scattering in the Parker spiral, started on
    a sector of the 1-au sphere (InitialSpaceSphereSector: theta in [1.2, 1.9], phi in [-0.3, 0.4])."""
    s = cf.Setup("gc_sector_start", cf.TRAJ_GUIDING_DIFF_SCATT, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    s.add(cf.DIFFUSION, "DiffusionPerpConstantPitchIsotropic", DC().insert_double(8.6e19 / cf.unit_diffusion).insert_double(0.02))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceSphereSector", DC().insert_vec((0.0, 0.0, 0.0)).insert_double(cf.one_au).insert_double(1.2)
          .insert_double(1.9).insert_double(-0.3).insert_double(0.4))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(30.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.02 * cf.one_day)))
    d = dist((16, 16, 0), (0.9 * cf.one_au, -1.0, 0.0), (1.2 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def soqlt(lam_par, lam_perp, R0, bmix_idx, kr_red, r_lim, sc_idx, sc_eff):
    return (DC().insert_double(lam_par).insert_double(lam_perp).insert_double(R0).insert_double(5.0e-6 / cf.unit_magnetic)
            .insert_int(bmix_idx).insert_double(kr_red).insert_double(r_lim).insert_int(sc_idx).insert_double(sc_eff))


def parker_soqlt_radial_limit():
    """Parker (backward) on the SolarWind background with DiffusionEmpiricalSOQLTandUNLT: the bent power laws
    on both sides of R0, the perpendicular mean free path reduced inside 3 au (Bmix_idx 1: region[1] = -1),
    the solar-cycle divisor; started at 3 au so that trajectories cross the reduction radius."""
    s = cf.Setup("parker_soqlt_radial_limit", cf.TRAJ_PARKER, cf.BACKWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    p0 = cf.mom_from_ekin(1.0 * cf.GeV)
    s.add(cf.DIFFUSION, "DiffusionEmpiricalSOQLTandUNLT", soqlt(0.03 * cf.one_au, 0.001 * cf.one_au, p0 * cf.c_code, 1, 0.3,
                                                                 3.0 * cf.one_au, 0, 0.25))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.6 * 3.0 * cf.one_au, 0.64 * 3.0 * cf.one_au, 0.48 * 3.0 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(0.3 * p0).insert_double(3.0 * p0).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 4.0 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", -20.0 * cf.one_day)))
    d = dist((8, 1, 1), (0.1 * p0, -1.0, -1.0), (10.0 * p0, 1.0, 1.0), (1, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(0)
    s.add(cf.DISTRIBUTION, "DistributionMomentumUniform", d)
    return s


def gc_soqlt_diff():
    """GuidingDiff in the Parker spiral with DiffusionEmpiricalSOQLTandUNLT: kappa_perp ~ 1/|B|, reduced inside
    1.25 au (Bmix_idx 1), and its perpendicular drift by DiffusionBase's directional derivative."""
    s = cf.Setup("gc_soqlt_diff", cf.TRAJ_GUIDING_DIFF, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundSolarWind", cf.solarwind())
    p0 = cf.mom_from_ekin(1.0 * cf.GeV)
    s.add(cf.DIFFUSION, "DiffusionEmpiricalSOQLTandUNLT", soqlt(2.0 * cf.one_au, 0.002 * cf.one_au, p0 * cf.c_code, 1, 0.4,
                                                                 1.25 * cf.one_au, 1, 0.3))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((1.2 * cf.one_au, 0.4 * cf.one_au, 0.1 * cf.one_au)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(5.0 * cf.MeV)).insert_double(cf.mom_from_ekin(500.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 1.4 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundarySphereAbsorb", cf.boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 0.3 * cf.one_au)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 0.1 * cf.one_day)))
    d = dist((16, 16, 0), (1.0 * cf.one_au, -1.0, 0.0), (1.5 * cf.one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionPositionPitchUniform", d)
    return s


def shock_oblique(u0=(0.002, 0.001, 0.0), B=(3.0, 4.0, 12.0), compression=3.0, v_shock=0.0):
    """BackgroundShock with its normal along -x through the origin: upstream is x < v_shock t (the reference's
    (x - r0) . n - v_shock t > 0), the downstream state from the jump conditions (src/background_shock.cc:58-86)."""
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), u0, B, 1.0)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(v_shock).insert_double(compression)
    return dc


def lorentz_shock():
    """TrajectoryLorentz (forward) across a discontinuous oblique shock: E = E0 upstream, E1 downstream (the
    background's own induced fields, src/background_shock.cc:91-111), B jumping at x = 0; isotropic momenta
    on a thick shell started just upstream, planes beyond either side and a time limit."""
    s = cf.Setup("lorentz_shock", cf.TRAJ_LORENTZ, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundShock", shock_oblique())
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((-0.05, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_double(cf.mom_from_ekin(100.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryPlaneReflect", cf.boundary(-1, [-1], ("vec", (0.0, 0.0, -3.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryPlanePass", cf.boundary(-1, [0], ("vec", (0.0, 0.0, 6.0)), ("vec", (0.0, 0.0, 1.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 40.0)))
    d = dist((10, 0, 0), (0.0, 0, 0), (44.0, 0, 0))
    d.insert_double(1.0).insert_double(0.5).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def focused_shock():
    """TrajectoryFocused (forward) across a moving discontinuous shock: U and B by side, every derivative
    zero (BackgroundShock::EvaluateBackgroundDerivatives, src/background_shock.cc:116-128); started upstream
    on a shell, a pass plane downstream, a time limit, binned in time."""
    s = cf.Setup("focused_shock", cf.TRAJ_FOCUSED, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundShock", shock_oblique(u0=(0.02, 0.0, 0.0), B=(2.0, 0.0, 1.0), v_shock=-0.005))
    p0 = cf.mom_from_ekin(10.0 * cf.MeV)
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((-0.2, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(p0))
    s.add(cf.BOUNDARY, "BoundaryPlanePass", cf.boundary(-1, [0], ("vec", (3.0, 0.0, 0.0)), ("vec", (1.0, 0.0, 0.0))))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 60.0)))
    d = dist((10, 0, 0), (0.0, 0, 0), (66.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def small_spiral_grid():
    """The Parker spiral sampled on 8^3 blocks of 4^3 zones over +-8 au (0.5 au zones)."""
    from spectrum_amd import cartesian
    return cartesian.parker_spiral(nb=8, corner_au=8.0)


def lorentz_grid():
    """TrajectoryLorentz (backward) on the gridded solar wind: B and |B| interpolated, E from the interpolated
    flow and field (the server's own), dmax limited by the zone; the lorentz_spiral_backward problem."""
    s = lorentz_spiral_backward()
    s.name, s.grid = "lorentz_grid", small_spiral_grid()
    s.plugins = [x for x in s.plugins if x[0] != cf.BACKGROUND]
    s.add(cf.BACKGROUND, "BackgroundServerCartesian", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), cf.one_au))
    return s


def focused_grid():
    """TrajectoryFocused (forward) on the gridded solar wind: the server's stencil gradients of U and B give
    the focusing, cooling and drift terms; the focused_spiral problem."""
    s = focused_spiral()
    s.name, s.grid = "focused_grid", small_spiral_grid()
    s.plugins = [x for x in s.plugins if x[0] != cf.BACKGROUND]
    s.add(cf.BACKGROUND, "BackgroundServerCartesian", cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), cf.one_au))
    return s


def gc_smoothshock_diffscatt():
    """GuidingDiffScatt across a moving oblique smooth shock: analytic grad B and grad|B|, the time-dependent
    terms of E* (d bhat / dt) and dp_perp/dt (d|B|/dt), dmax from the shock coordinate of the previous
    evaluation carried per slot; isotropic pitch-angle scattering and perpendicular diffusion."""
    s = cf.Setup("gc_smoothshock_diffscatt", cf.TRAJ_GUIDING_DIFF_SCATT, cf.FORWARD)
    w = 0.002 * cf.one_au
    B = 5.0e-5 / cf.unit_magnetic
    dc = cf.background_base(DC(), 0.0, (0.0, 0.0, 0.0), (0.012, 0.003, 0.0), (B, 2.0 * B, 0.5 * B), 10.0 * w)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(-0.004).insert_double(3.0).insert_double(w).insert_double(0.1)
    s.add(cf.BACKGROUND, "BackgroundSmoothShock", dc)
    s.add(cf.DIFFUSION, "DiffusionPerpConstantPitchIsotropic", DC().insert_double(2.0e17 / cf.unit_diffusion).insert_double(0.05))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((0.15 * w, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumShell", DC().insert_double(cf.mom_from_ekin(2.0 * cf.MeV)))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 30.0)))
    d = dist((10, 0, 0), (0.0, 0, 0), (60.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


def guiding_shock():
    """TrajectoryGuiding across the discontinuous oblique shock: E0 / E1 and B by side, no gradients; started
    upstream on a thick shell (isotropic mu), a time limit."""
    s = cf.Setup("guiding_shock", cf.TRAJ_GUIDING, cf.FORWARD)
    s.add(cf.BACKGROUND, "BackgroundShock", shock_oblique(u0=(0.02, 0.004, 0.0), B=(1.0, 2.0, 0.5), v_shock=-0.002))
    s.add(cf.INITIAL, "InitialTimeFixed", DC().insert_double(0.0))
    s.add(cf.INITIAL, "InitialSpaceFixed", DC().insert_vec((-0.5, 0.0, 0.0)))
    s.add(cf.INITIAL, "InitialMomentumThickShell", DC().insert_double(cf.mom_from_ekin(1.0 * cf.MeV)).insert_double(cf.mom_from_ekin(20.0 * cf.MeV)).insert_bool(True))
    s.add(cf.BOUNDARY, "BoundaryTimeExpire", cf.boundary(1, [0], ("double", 60.0)))
    d = dist((10, 0, 0), (0.0, 0, 0), (120.0, 0, 0), (0, 0, 0), (1, 0, 0))
    d.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(cf.DISTRIBUTION, "DistributionTimeUniform", d)
    return s


ALL = [shock_flowmom_reflect, solarwind_kinrad_reflect_recurrent, smoothshock_perp_momentum, uniform_para_plane_pass,
       lambda: modulation_spectrum_type(0), lambda: modulation_spectrum_type(2), gc_rigidity_reflect, gc_isotropic_only,
       guiding_spiral, guiding_scatt_uniform, guiding_diff_rigidity, parker_strauss_heliopause, parker_potgieter_uniform,
       gc_qlt_spiral, fieldline_spiral, fieldline_backward_shock, lorentz_uniform, lorentz_spiral_backward,
       focused_spiral, focused_uniform_backward, gc_wnlt_scatt_spiral, gc_wnlt_ramp_diffscatt,
       gc_strauss_diff, focused_smoothshock, uniform_sphere_start, gc_sector_start, parker_soqlt_radial_limit, gc_soqlt_diff,
       lorentz_shock, focused_shock, lorentz_grid, focused_grid, gc_smoothshock_diffscatt, guiding_shock, parker_offset_spheres]
NAMES = ["shock_flowmom_reflect", "solarwind_kinrad_reflect_recurrent", "smoothshock_perp_momentum", "uniform_para_plane_pass",
         "modulation_type0", "modulation_type2", "gc_rigidity_reflect", "gc_isotropic_only",
         "guiding_spiral", "guiding_scatt_uniform", "guiding_diff_rigidity", "parker_strauss_heliopause", "parker_potgieter_uniform",
         "gc_qlt_spiral", "fieldline_spiral", "fieldline_backward_shock", "lorentz_uniform", "lorentz_spiral_backward",
         "focused_spiral", "focused_uniform_backward", "gc_wnlt_scatt_spiral", "gc_wnlt_ramp_diffscatt",
         "gc_strauss_diff", "focused_smoothshock", "uniform_sphere_start", "gc_sector_start", "parker_soqlt_radial_limit",
         "gc_soqlt_diff", "lorentz_shock", "focused_shock", "lorentz_grid", "focused_grid", "gc_smoothshock_diffscatt",
         "guiding_shock", "parker_offset_spheres"]


# Setups whose endpoints amplify last-bit differences far beyond 1e-10: QLT's mu derivative is
# DiffusionBase::GetMuDerivative's forward difference with dmu = 1e-8 (src/diffusion_base.cc:238), which
# multiplies rounding differences by ~1e8 (the WNLT models inherit it).  Their parity bound is scaled by the oracle's own disagreement
# with its last-bit-perturbed build (test_gpu_parity.sensitivity).  The focused transport along the spiral
# differentiates the field over the Larmor radius (a difference of nearby field values divided by ~1e-3 of
# the distance to the Sun), and its long trajectories reach a few 1e-10 of last-bit sensitivity.
ILL_CONDITIONED = {"gc_qlt_spiral", "focused_spiral", "gc_wnlt_scatt_spiral", "gc_wnlt_ramp_diffscatt", "focused_smoothshock"}


def weight_atol(name, rec, rtol):
    """Absolute bound on the difference of recorded event weights between two runs whose endpoints
    agree to `rtol` relative: the <dx dx> weight x_i^2 moves by at most 2 |x| |dx| <= 2 rtol |x|^2."""
    atol = np.zeros((rec.shape[0], 4))
    if name == "smoothshock_perp_momentum":
        atol[:, 0] = 2.0 * rtol * np.sum((rec["pos"] - rec["pos0"]) ** 2, axis=1) * (1.0 + 1e-6)
    return atol
