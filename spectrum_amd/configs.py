"""Code units and the plugin setups of the measured workloads.

Units follow common/specie.hh:30-85 of the reference (length 1e10 cm, velocity
0.1 c, proton mass) with its in-tree constant values
(common/physical_constants.hh:20-81).  Each builder returns a ``Setup``: the
transport, the time flow, runtime switches, and the ordered plugin list
``(role, class name, DataContainer bytes)`` that both the HIP path and the CPU
oracle consume.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from .data_container import DataContainer

# --- constants (common/physical_constants.hh, USE_GSL off) -------------------------------
C_CGS = 2.99792458e+10
MP_CGS = 1.67262192596e-24
E_CGS = 4.8032044e-10
EV_CGS = 1.602176634e-12
AU_CGS = 1.495978707e+13

unit_length = 1.0e10
unit_velocity = 0.1 * C_CGS
unit_mass = MP_CGS
unit_charge = E_CGS
unit_momentum = unit_mass * unit_velocity
unit_energy = unit_mass * unit_velocity * unit_velocity
unit_time = unit_length / unit_velocity
unit_magnetic = unit_momentum / (unit_charge * unit_time)
unit_rigidity = unit_energy / unit_charge
unit_diffusion = unit_velocity * unit_length
c_code = C_CGS / unit_velocity
mass = MP_CGS / unit_mass
charge = E_CGS / unit_charge

one_au = AU_CGS / unit_length
one_day = 24.0 * 60.0 * 60.0 / unit_time
one_year = 60.0 * 60.0 * 24.0 * 365.0 / unit_time
MeV = 1.0e6 * EV_CGS / unit_energy
GeV = 1.0e9 * EV_CGS / unit_energy
keV = 1.0e3 * EV_CGS / unit_energy
GV = 1.0e9 / 299.792458 / unit_rigidity   # 1 GV in statvolt, in code units


def mom_from_ekin(T: float) -> float:
    """Particle::Mom (common/physics.hh:262-266)."""
    return math.sqrt(T * (T + 2.0 * mass * c_code * c_code)) / c_code


def ekin_from_mom(p: float) -> float:
    """Particle::EnrKin (common/physics.hh:278-282)."""
    return math.sqrt(p * p + (mass * c_code) ** 2) * c_code - mass * c_code * c_code


def vel_from_mom(p: float) -> float:
    return p * c_code * c_code / (math.sqrt(p * p + (mass * c_code) ** 2) * c_code)


# roles of the plugin families
BACKGROUND, DIFFUSION, SOURCE, BOUNDARY, INITIAL, DISTRIBUTION = range(6)
ROLE_NAMES = ["background", "diffusion", "source", "boundary", "initial", "distribution"]

# transports and time flows
TRAJ_PARKER, TRAJ_PARKER_SOURCE, TRAJ_GUIDING_DIFF_SCATT, TRAJ_GUIDING, TRAJ_GUIDING_SCATT, TRAJ_GUIDING_DIFF, TRAJ_FIELDLINE, TRAJ_LORENTZ, TRAJ_FOCUSED = 0, 1, 2, 3, 4, 5, 6, 7, 8
FORWARD, BACKWARD = 0, 1


@dataclass
class Setup:
    name: str
    traj_type: int
    time_flow: int
    plugins: list = field(default_factory=list)     # (role, class name, bytes)
    options: dict = field(default_factory=dict)     # runtime replacements of compile-time switches
    meta: dict = field(default_factory=dict)
    grid: object = None                             # cartesian.CartesianGrid of a BackgroundServerCartesian

    def add(self, role: int, cls: str, dc: DataContainer) -> "Setup":
        self.plugins.append((role, cls, dc.bytes()))
        return self


# --- plugin container builders (parameter order = the reference's Setup* Read order) ------

def background_base(dc: DataContainer, t0=0.0, r0=(0.0, 0.0, 0.0), u0=(0.0, 0.0, 0.0), B0=(0.0, 0.0, 0.0), dmax0=1.0):
    """BackgroundBase::SetupBackground (src/background_base.cc:334-352)."""
    dc.insert_double(t0).insert_vec(r0).insert_vec(u0).insert_vec(B0).insert_double(dmax0)
    return dc


def solarwind(V_kms=400.0, B_nT_1au=5.0, omega=2.87e-6, r_ref=3.0 * 6.957e10 / unit_length, dmax0=None, dmax_fraction=0.1,
              axis=(0.0, 0.0, 1.0)):
    """BackgroundSolarWind parameters (src/background_solarwind.hh:87, src/background_solarwind.cc:58-84); the
    rotation vector is omega along `axis`."""
    umag = V_kms * 1.0e5 / unit_velocity
    BmagE = B_nT_1au * 1.0e-5 / unit_magnetic
    Bref = BmagE * (one_au / r_ref) ** 2
    dc = background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), (umag, 0.0, 0.0), (Bref, 0.0, 0.0), one_au if dmax0 is None else dmax0)
    na = math.sqrt(sum(float(v) * float(v) for v in axis))
    dc.insert_vec(tuple(omega * unit_time * float(v) / na for v in axis)).insert_double(r_ref).insert_double(dmax_fraction)
    return dc


def boundary(max_crossings: int, actions, *extra) -> DataContainer:
    dc = DataContainer().insert_int(max_crossings).insert_int_vector(actions)
    for kind, val in extra:
        getattr(dc, "insert_" + kind)(val)
    return dc


def distribution_base(dc: DataContainer, n_bins, minval, maxval, log_bins, bin_outside, unit_distro=1.0, unit_val=(1.0, 1.0, 1.0), keep_records=False):
    """DistributionTemplated::SetupDistribution (src/distribution_templated.cc:92-110)."""
    dc.insert_multiindex(n_bins).insert_vec(minval).insert_vec(maxval).insert_multiindex(log_bins).insert_multiindex(bin_outside)
    dc.insert_double(unit_distro).insert_vec(unit_val).insert_bool(keep_records)
    return dc


def energy_grid(T1: float, T2: float, n: int) -> list:
    """Log-spaced kinetic energies of InitialMomentumShellGrid."""
    if n == 1:
        return [T1]
    l1, l2 = math.log10(T1), math.log10(T2)
    return [10.0 ** (l1 + (l2 - l1) * k / (n - 1)) for k in range(n)]


def spectrum_bins(T1: float, T2: float, n: int):
    """Log bins centred on each grid energy, so energy k lands in bin k."""
    if n == 1:
        return T1 / 1.1, T1 * 1.1
    half = 0.5 * (math.log10(T2) - math.log10(T1)) / (n - 1)
    return 10.0 ** (math.log10(T1) - half), 10.0 ** (math.log10(T2) + half)


# --- workloads --------------------------------------------------------------------------

def parker_modulation(n_energies=16, T1=0.1 * GeV, T2=10.0 * GeV, r_start_au=1.0, r_outer_au=90.0,
                      t_limit_years=50.0, spectrum_type=1, inner_au=0.01, time_action=-1,
                      name="parker_modulation", sw=None) -> Setup:
    """Backward Parker modulation in the Parker spiral (the CPU reference problem and its GPU scale-up).

    SolarWind V = 400 km/s, |B| = 5 nT at 1 au, Omega = 2.87e-6 rad/s (r_ref = 3 R_sun as in
    runs/gcr_modulation_example.cc:52-57); kappa_par = 1.5e22 cm^2/s * beta * (R / 1 GV)^0.5,
    kappa_perp = 0.02 kappa_par; start at r = 1 au on the ecliptic at n log-spaced kinetic
    energies; absorbing sphere at 90 au weighted by a LIS j ~ T^-2.7 (intensity weighting).
    An inner absorbing sphere and a time limit without binning follow the reference example
    (``inner_au=None`` drops the sphere; ``time_action=0`` bins trajectories at the time limit too;
    ``sw`` replaces the solar wind's parameters, e.g. ``solarwind(axis=...)``).
    """
    s = Setup(name, TRAJ_PARKER, BACKWARD, options={"DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE": spectrum_type})
    s.add(BACKGROUND, "BackgroundSolarWind", solarwind() if sw is None else sw)
    lam0 = 3.0 * (1.5e22 / unit_diffusion) / c_code     # kappa_par = lam0 * v / 3 with v = beta c
    dif = DataContainer().insert_double(lam0).insert_double(GV).insert_double(5.0e-5 / unit_magnetic)
    dif.insert_double(0.5).insert_double(0.0).insert_double(0.02)
    s.add(DIFFUSION, "DiffusionRigidityMagneticFieldPowerLaw", dif)
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((r_start_au * one_au, 0.0, 0.0)))
    s.add(INITIAL, "InitialMomentumShellGrid", DataContainer().insert_double(T1).insert_double(T2).insert_int(n_energies))
    if inner_au is not None:
        s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [-1], ("vec", (0.0, 0.0, 0.0)), ("double", inner_au * one_au)))
    s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", r_outer_au * one_au)))
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [time_action], ("double", -t_limit_years * one_year)))
    lo, hi = spectrum_bins(T1, T2, n_energies)
    dst = distribution_base(DataContainer(), (n_energies, 0, 0), (lo, 0.0, 0.0), (hi, 0.0, 0.0), (1, 0, 0), (0, 0, 0),
                            1.0, (unit_energy, 1.0, 1.0), False)
    dst.insert_double(1.0).insert_double(GeV).insert_double(-2.7).insert_double(0.0)
    s.add(DISTRIBUTION, "DistributionSpectrumKineticEnergyPowerLaw", dst)
    s.meta.update(energies=energy_grid(T1, T2, n_energies), n_energies=n_energies)
    return s


def dsa_shock(n_bins=128, p_ratio=1.0e4, downstream_kappa_lengths=4.0, t_limit_days=3000.0, src_ratio=1.0,
              name="dsa_smooth_shock") -> Setup:
    """Diffusive shock acceleration at a planar smooth shock, forward ParkerSource.

    u1 = 400 km/s, r = 4, width = 0.01 kappa / u1, kappa = 1e20 cm^2/s, 10 keV protons injected at the
    shock; a downstream absorbing plane at a few kappa/u2 and a long time limit; momentum at escape
    binned on n_bins log bins over [p_inj, p_ratio p_inj].  Expected counts per log bin ~ p^(3-q), q = 3r/(r-1).
    The source term (SourceMomentumInjection) fires when the momentum crosses src_ratio * p_inj; with
    the default ratio 1 the particles start on it and it never fires.
    """
    s = Setup(name, TRAJ_PARKER_SOURCE, FORWARD)
    u1 = 4.0e7 / unit_velocity
    kappa = 1.0e20 / unit_diffusion
    width = 0.01 * kappa / u1
    B = 5.0e-5 / unit_magnetic
    dc = background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), (u1, 0.0, 0.0), (B, 0.0, 0.0), 10.0 * width)
    dc.insert_vec((-1.0, 0.0, 0.0)).insert_double(0.0).insert_double(4.0).insert_double(width).insert_double(0.1)
    s.add(BACKGROUND, "BackgroundSmoothShock", dc)
    s.add(DIFFUSION, "DiffusionFullConstant", DataContainer().insert_double(kappa).insert_double(kappa))
    p0 = mom_from_ekin(10.0 * keV)
    s.add(SOURCE, "SourceMomentumInjection", DataContainer().insert_double(src_ratio * p0).insert_double(1.0))
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((0.0, 0.0, 0.0)))
    s.add(INITIAL, "InitialMomentumShell", DataContainer().insert_double(p0))
    L_down = downstream_kappa_lengths * kappa / (u1 / 4.0)
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [-1], ("double", t_limit_days * one_day)))
    s.add(BOUNDARY, "BoundaryPlaneAbsorb", boundary(1, [0], ("vec", (L_down, 0.0, 0.0)), ("vec", (1.0, 0.0, 0.0))))
    dst = distribution_base(DataContainer(), (n_bins, 1, 1), (p0, -1.0, -1.0), (p_ratio * p0, 1.0, 1.0), (1, 0, 0), (0, 0, 0))
    dst.insert_double(1.0).insert_double(0.0).insert_int(1).insert_int(0)
    s.add(DISTRIBUTION, "DistributionMomentumUniform", dst)
    s.meta.update(p0=p0, kappa=kappa, u1=u1, width=width, compression=4.0, L_down=L_down)
    return s


def gc_diff_scatt(n_r=32, n_mu=64, T_MeV=100.0, nu=1.0e-4, kperp_cgs=8.6e19, t_limit_days=30.0, name="gc_diff_scatt", sw=None) -> Setup:
    """Guiding centre with perpendicular diffusion and pitch-angle scattering in the Parker spiral.

    D_mumu = nu (1 - mu^2) / 2 with nu = 1e-4 1/s; kappa_perp constant (the value is this build's choice:
    0.02 kappa_par of 100 MeV protons in the C1 model); 100 MeV protons with isotropic mu from 1 au;
    a 30-day time boundary; (r, mu) distribution at expiry.  ``sw`` replaces the solar wind's parameters.
    """
    s = Setup(name, TRAJ_GUIDING_DIFF_SCATT, FORWARD)
    s.add(BACKGROUND, "BackgroundSolarWind", solarwind() if sw is None else sw)
    D0 = 0.5 * nu * unit_time
    s.add(DIFFUSION, "DiffusionPerpConstantPitchIsotropic", DataContainer().insert_double(kperp_cgs / unit_diffusion).insert_double(D0))
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((one_au, 0.0, 0.0)))
    s.add(INITIAL, "InitialMomentumShell", DataContainer().insert_double(mom_from_ekin(T_MeV * MeV)))
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [0], ("double", t_limit_days * one_day)))
    dst = distribution_base(DataContainer(), (n_r, n_mu, 0), (0.0, -1.0, 0.0), (3.0 * one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    dst.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(DISTRIBUTION, "DistributionPositionPitchUniform", dst)
    s.meta.update(D0=D0, p=mom_from_ekin(T_MeV * MeV))
    return s


def gc_uniform(D0=1.0e-3, kperp=0.5, T_MeV=100.0, theta0=None, t_limit=500.0, B=(0.0, 0.0, 10.0), name="gc_uniform") -> Setup:
    """Guiding centre in a uniform field with U = 0: free streaming along B, pitch-angle scattering
    D_mumu = D0 (1 - mu^2) and constant perpendicular diffusion; started at the origin either as a
    ring at pitch angle theta0 (InitialMomentumRing) or isotropic (InitialMomentumShell), stopped and
    binned at t_limit.  Closed forms: <P_l(mu)> decays as exp(-l(l+1) D0 t), <x^2 + y^2> = 4 kperp t,
    and for an isotropic start <z^2> = (2 v^2 / 3) [t / (2 D0) - (1 - exp(-2 D0 t)) / (4 D0^2)]."""
    s = Setup(name, TRAJ_GUIDING_DIFF_SCATT, FORWARD)
    s.add(BACKGROUND, "BackgroundUniform", background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), B, 1.0e6))
    s.add(DIFFUSION, "DiffusionPerpConstantPitchIsotropic", DataContainer().insert_double(kperp).insert_double(D0))
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((0.0, 0.0, 0.0)))
    p = mom_from_ekin(T_MeV * MeV)
    if theta0 is None:
        s.add(INITIAL, "InitialMomentumShell", DataContainer().insert_double(p))
    else:
        s.add(INITIAL, "InitialMomentumRing", DataContainer().insert_double(p).insert_double(theta0))
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [0], ("double", t_limit)))
    dst = distribution_base(DataContainer(), (10, 0, 0), (0.0, 0, 0), (1.1 * t_limit, 0, 0), (0, 0, 0), (0, 0, 0))
    dst.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(DISTRIBUTION, "DistributionTimeUniform", dst)
    s.meta.update(D0=D0, kperp=kperp, p=p, v=vel_from_mom(p), theta0=theta0, t_limit=t_limit)
    return s


def uniform_walk(D=2.0, U=(0.0, 0.0, 0.0), B=(3.0, 4.0, 0.0), t_limit=1000.0, dmax=5.0, flow=BACKWARD, name="uniform_walk") -> Setup:
    """Parker transport in a uniform background with constant isotropic diffusion (DiffusionFullConstant
    with kappa_perp = kappa_par = D): x(t) - x0 - (-/+) U t is Gaussian with variance 2 D t per axis."""
    s = Setup(name, TRAJ_PARKER, flow)
    s.add(BACKGROUND, "BackgroundUniform", background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), U, B, dmax))
    s.add(DIFFUSION, "DiffusionFullConstant", DataContainer().insert_double(D).insert_double(D))
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((1.0, 2.0, 3.0)))
    s.add(INITIAL, "InitialMomentumShell", DataContainer().insert_double(mom_from_ekin(MeV)))
    tmark = -t_limit if flow == BACKWARD else t_limit
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [0], ("double", tmark)))
    lo, hi = (tmark * 1.1, 0.0) if flow == BACKWARD else (0.0, tmark * 1.1)
    dst = distribution_base(DataContainer(), (10, 0, 0), (lo, 0, 0), (hi, 0, 0), (0, 0, 0), (0, 0, 0))
    dst.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(DISTRIBUTION, "DistributionTimeUniform", dst)
    s.meta.update(D=D, U=U, t_limit=t_limit)
    return s


def parker_modulation_cartesian(grid, n_bins=100, T1_MeV=10.0, T2_MeV=5000.0, inner_au=0.05, outer_au=80.0,
                                t_limit_years=1.0, kap_rat=1.0, name="parker_modulation_cartesian") -> Setup:
    """Backward Parker modulation on a gridded solar wind, benchmarks/main_test_modulation_cartesian_parker.cc:
    BackgroundServerCartesian (dmax 1 au) on `grid` (cartesian.parker_spiral() is the reference's
    25^3-block background), kappa = 1.5e22 cm^2/s (T / 1 GeV) (r / 1 au)^2 isotropic, log-uniform
    10 MeV - 5 GeV from 1 au, absorbing spheres at 0.05 au (action 1) and 80 au (the hot boundary,
    action 0), a 1-year limit (not binned), and a 100-bin spectrum with J0 T^-1.8 at the outer boundary.
    The unused shock parameters of the diffusion model go in as ints, as the benchmark inserts them.
    ``kap_rat`` != 1 makes kappa anisotropic (the benchmark's is isotropic)."""
    s = Setup(name, TRAJ_PARKER, BACKWARD, grid=grid)
    s.add(BACKGROUND, "BackgroundServerCartesian", background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), one_au))
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((one_au, 0.0, 0.0)))
    p1, p2 = mom_from_ekin(T1_MeV * MeV), mom_from_ekin(T2_MeV * MeV)
    s.add(INITIAL, "InitialMomentumThickShell", DataContainer().insert_double(p1).insert_double(p2).insert_bool(True))
    s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [1], ("vec", (0.0, 0.0, 0.0)), ("double", inner_au * one_au)))
    s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", outer_au * one_au)))
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [-1], ("double", -t_limit_years * one_year)))
    dif = DataContainer().insert_double(1.5e22 / unit_diffusion).insert_double(GeV).insert_double(one_au)
    dif.insert_double(1.0).insert_double(2.0).insert_double(kap_rat).insert_int(0)
    dif.insert_int(1).insert_int(0).insert_int(1)
    s.add(DIFFUSION, "DiffusionKineticEnergyRadialDistancePowerLaw", dif)
    unit_distro = 1.0 / (unit_length ** 2 * unit_time * 4.0 * math.pi * unit_energy)
    dst = distribution_base(DataContainer(), (n_bins, 0, 0), (T1_MeV * MeV, 0.0, 0.0), (T2_MeV * MeV, 0.0, 0.0), (1, 0, 0), (0, 0, 0),
                            unit_distro, (unit_energy, 1.0, 1.0), False)
    dst.insert_double(1.0 / unit_distro).insert_double(GeV).insert_double(-1.8).insert_double(0.0)
    s.add(DISTRIBUTION, "DistributionSpectrumKineticEnergyPowerLaw", dst)
    # grid interpolations per step: the new position, plus the six derivative points of div(kappa)
    # unless kappa is isotropic and reads no field there (then only their positions are checked)
    s.meta.update(T1=T1_MeV * MeV, T2=T2_MeV * MeV, n_bins=n_bins, grid_gathers_per_step=1 if kap_rat == 1.0 else 7)
    return s


def parker_modulation_radial(n_energies=8, T1_MeV=10.0, T2_MeV=5000.0, inner_au=0.05, outer_au=80.0, t_limit_years=100.0,
                             spectrum_type=0, name="parker_modulation_radial", grid=None, grid_dmax_au=1.0) -> Setup:
    """The physics of the reference's Cartesian modulation benchmark (benchmarks/main_test_modulation_cartesian_parker.cc)
    on the analytic solar wind instead of its sampled grid: backward Parker, radial wind V = 400 km/s,
    isotropic kappa = 1.5e22 cm^2/s (T / 1 GeV) (r / 1 au)^2, absorbing spheres at 0.05 au (cold, 0) and
    80 au (hot, J0 [v] (T / 1 GeV)^-1.8 [/ p^2]), from 1 au at n log-spaced kinetic energies, and a time limit
    long enough that no trajectory reaches it (not binned).  With kappa isotropic the magnetic field drops
    out, so the spectrum at 1 au is the solution of the spherically symmetric Parker equation.  With ``grid``
    the wind is that grid's (BackgroundServerCartesian, dmax0 = ``grid_dmax_au``), e.g. the spiral sampled at
    zone centres (cartesian.parker_spiral), whose interpolated wind approaches the radial one as the zones shrink."""
    s = Setup(name, TRAJ_PARKER, BACKWARD, options={"DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE": spectrum_type}, grid=grid)
    if grid is None:
        s.add(BACKGROUND, "BackgroundSolarWind", solarwind())
    else:
        s.add(BACKGROUND, "BackgroundServerCartesian", background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0),
                                                                       grid_dmax_au * one_au))
    dif = DataContainer().insert_double(1.5e22 / unit_diffusion).insert_double(GeV).insert_double(one_au)
    dif.insert_double(1.0).insert_double(2.0).insert_double(1.0).insert_int(0)
    dif.insert_double(1.0).insert_double(0.0).insert_double(1.0)
    s.add(DIFFUSION, "DiffusionKineticEnergyRadialDistancePowerLaw", dif)
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((one_au, 0.0, 0.0)))
    T1, T2 = T1_MeV * MeV, T2_MeV * MeV
    s.add(INITIAL, "InitialMomentumShellGrid", DataContainer().insert_double(T1).insert_double(T2).insert_int(n_energies))
    s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [1], ("vec", (0.0, 0.0, 0.0)), ("double", inner_au * one_au)))
    s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", outer_au * one_au)))
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [-1], ("double", -t_limit_years * one_year)))
    lo, hi = spectrum_bins(T1, T2, n_energies)
    dst = distribution_base(DataContainer(), (n_energies, 0, 0), (lo, 0.0, 0.0), (hi, 0.0, 0.0), (1, 0, 0), (0, 0, 0),
                            1.0, (unit_energy, 1.0, 1.0), False)
    dst.insert_double(1.0).insert_double(GeV).insert_double(-1.8).insert_double(0.0)
    s.add(DISTRIBUTION, "DistributionSpectrumKineticEnergyPowerLaw", dst)
    s.meta.update(energies=energy_grid(T1, T2, n_energies), n_energies=n_energies, inner=inner_au * one_au, outer=outer_au * one_au,
                  V=400.0e5 / unit_velocity, kap0=1.5e22 / unit_diffusion, T0=GeV, r0=one_au, gamma=-1.8)
    return s


def gc_cartesian(grid, T_MeV=100.0, D0=1.668e-4, kperp_cgs=8.6e19, t_limit_days=2.0, r_start_au=1.0, name="gc_cartesian") -> Setup:
    """Guiding centre with perpendicular diffusion and pitch-angle scattering on a gridded solar wind
    (BackgroundServerCartesian, dmax 1 au): the gc_diff_scatt workload with the analytic spiral replaced
    by its sampled grid; (r, mu) map at the time limit."""
    s = Setup(name, TRAJ_GUIDING_DIFF_SCATT, FORWARD, grid=grid)
    s.add(BACKGROUND, "BackgroundServerCartesian", background_base(DataContainer(), 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), one_au))
    s.add(DIFFUSION, "DiffusionPerpConstantPitchIsotropic", DataContainer().insert_double(kperp_cgs / unit_diffusion).insert_double(D0))
    s.add(INITIAL, "InitialTimeFixed", DataContainer().insert_double(0.0))
    s.add(INITIAL, "InitialSpaceFixed", DataContainer().insert_vec((r_start_au * one_au, 0.0, 0.0)))
    s.add(INITIAL, "InitialMomentumShell", DataContainer().insert_double(mom_from_ekin(T_MeV * MeV)))
    s.add(BOUNDARY, "BoundarySphereAbsorb", boundary(1, [0], ("vec", (0.0, 0.0, 0.0)), ("double", 60.0 * one_au)))
    s.add(BOUNDARY, "BoundaryTimeExpire", boundary(1, [0], ("double", t_limit_days * one_day)))
    dst = distribution_base(DataContainer(), (32, 64, 0), (0.0, -1.0, 0.0), (10.0 * one_au, 1.0, 1.0), (0, 0, 0), (1, 0, 0))
    dst.insert_double(1.0).insert_double(0.0).insert_int(1)
    s.add(DISTRIBUTION, "DistributionPositionPitchUniform", dst)
    return s
