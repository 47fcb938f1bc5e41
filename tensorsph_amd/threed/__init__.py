"""3-D operators, periodic search, conservation ledger and the shear-wave
driver of tensorsph.threed (threed/__init__.py:31-91), on libtsph.

Not built here (outside the particle pipeline): the verification helpers
laplacian_of_field / gradient_of_field / the moment diagnostics,
run_hydrostatic, wall_shell, the surface-tension modules.
"""
from .conservation import (angular_momentum, centre_of_mass, conservation_row, diagnostics,
                           internal_energy, kinetic_energy, mass_total, momentum, relative_drift,
                           total_energy)
from .kernel3 import (SUPPORT_FACTOR, calibrated_particle_mass, cubic_spline_gradient,
                      cubic_spline_kernel, cubic_spline_laplacian, lattice_bias, lattice_kernel_sum,
                      pair_kernels, sigma3, wendland_c2_gradient, wendland_c2_kernel)
from .neighbors3 import box_pairs, brute_force_pairs_min_image, minimum_image, periodic_pairs
from .operators3 import (accelerations, analytic_mode_decay, cell_dx, density_diffusion_rate_3d,
                         mode_amplitude, pressure_acceleration, shear_wave_field, summation_density,
                         tait_density, tait_pressure, viscous_acceleration, wall_pressure_adami,
                         wavenumber_for)
from .solver import (PeriodicSimulation, ShearWaveResult, acoustic_dt, cubic_lattice, fit_decay_rate,
                     run_shear_wave, shear_wave_velocity)

__all__ = [
    "SUPPORT_FACTOR", "calibrated_particle_mass", "cubic_spline_gradient", "cubic_spline_kernel",
    "cubic_spline_laplacian", "wendland_c2_gradient", "wendland_c2_kernel", "lattice_bias",
    "lattice_kernel_sum", "pair_kernels", "sigma3", "box_pairs", "brute_force_pairs_min_image",
    "minimum_image", "periodic_pairs", "accelerations", "analytic_mode_decay", "cell_dx",
    "density_diffusion_rate_3d", "mode_amplitude", "pressure_acceleration", "shear_wave_field",
    "summation_density", "tait_density", "tait_pressure", "viscous_acceleration",
    "wall_pressure_adami", "wavenumber_for", "PeriodicSimulation", "ShearWaveResult", "acoustic_dt",
    "cubic_lattice", "fit_decay_rate", "run_shear_wave", "shear_wave_velocity", "angular_momentum",
    "centre_of_mass", "conservation_row", "diagnostics", "internal_energy", "kinetic_energy",
    "mass_total", "momentum", "relative_drift", "total_energy",
]
