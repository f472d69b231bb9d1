"""MI355X-native FFT-accelerated slab Foldy-Lax solver.

Drop-in for the slab path of tmnestor/MultipleScatteringCalculations
(cubic_scattering/slab_scattering.py, slab_scattering_gpu.py, torch_gmres.py,
solver_config.py): same public names and signatures, with every numerical step
in the HIP library libmsc_slab.so (include/msc_slab.h).
"""

from .effective_contrasts import (
    CubeTMatrixResult,
    MaterialContrast,
    ReferenceMedium,
    compute_cube_tmatrix,
    effective_stiffness_voigt,
)
from .slab_scattering import (
    SlabGeometry,
    SlabMaterial,
    SlabPlan,
    SlabResult,
    compute_slab_scattering,
    compute_slab_tmatrices,
    kennett_reference_rpp,
    random_iid_slab_material,
    random_slab_material,
    slab_reflected_field,
    slab_rpp_periodic,
    uniform_slab_material,
)
from .cpa_iteration import (
    CPAResult,
    CubicEffectiveMedium,
    Phase,
    compute_cpa,
    compute_cpa_two_phase,
    phases_from_two_phase,
    voigt_average,
)
from .kennett_layers import (
    FluidLayer,
    IsotropicLayer,
    KennettResult,
    LayerStack,
    PSVCoefficients,
    SHCoefficients,
    cpa_stack_from_phases,
    cubic_to_isotropic_layer,
    kennett_layers,
    psv_fluid_solid,
    random_heterogeneous_stack,
    random_velocity_stack,
)
from .learned_precond import LearnedPreconditioner, harvest_pairs, merge_pairs
from .ocean_bottom import (
    OceanBottomConfig,
    OceanBottomResult,
    compute_ocean_bottom_reflection,
    load_ocean_bottom_config,
    ricker_source_spectrum,
    write_log,
)
from .resonance_tmatrix import (
    ResonanceTmatrixResult,
    compute_resonance_tmatrices,
    compute_resonance_tmatrix,
    scattering_order_decomposition,
    strain_from_displacement_traction,
    sub_cell_centres,
    suggest_n_subcells,
    validate_rayleigh_limit,
    voigt_tmatrix_from_resonance_result,
)
from .resonance_slab import (
    ResonanceOceanBottomResult,
    compute_resonance_slab_scattering,
    compute_subcell_slab_scattering,
    resonance_ocean_bottom,
    resonance_slab_tmatrices,
    subcell_slab,
    uniform_contrast_phases,
)
from .slab_scattering_gpu import SlabOperator, compute_slab_scattering_gpu
from .sphere_scattering_fft import (
    SphereDecompositionResult,
    compute_sphere_foldy_lax,
    compute_sphere_foldy_lax_fft,
    compute_sphere_foldy_lax_fft_gpu,
)
from .solver_config import ScatteringConfig, load_config, run_from_config, validate_config
from .sweep import SweepCase, compute_slab_sweep
from .torch_gmres import get_device, select_dtype, to_numpy, to_torch, torch_gmres

__all__ = [
    "CubeTMatrixResult", "MaterialContrast", "ReferenceMedium", "compute_cube_tmatrix",
    "effective_stiffness_voigt", "SlabGeometry", "SlabMaterial", "SlabPlan", "SlabResult",
    "compute_slab_scattering", "compute_slab_scattering_gpu", "compute_slab_tmatrices",
    "random_iid_slab_material", "random_slab_material", "slab_reflected_field",
    "slab_rpp_periodic", "kennett_reference_rpp", "uniform_slab_material", "SlabOperator",
    "ScatteringConfig", "load_config", "run_from_config", "validate_config", "get_device", "select_dtype", "to_numpy", "to_torch", "torch_gmres",
    "LearnedPreconditioner", "harvest_pairs", "merge_pairs", "SweepCase", "compute_slab_sweep",
    "FluidLayer", "IsotropicLayer", "KennettResult", "LayerStack", "PSVCoefficients",
    "SHCoefficients", "kennett_layers", "psv_fluid_solid", "cubic_to_isotropic_layer",
    "cpa_stack_from_phases", "random_heterogeneous_stack", "random_velocity_stack",
    "OceanBottomConfig", "OceanBottomResult", "compute_ocean_bottom_reflection",
    "load_ocean_bottom_config", "ricker_source_spectrum", "write_log",
    "ResonanceTmatrixResult", "compute_resonance_tmatrices", "compute_resonance_tmatrix",
    "suggest_n_subcells", "validate_rayleigh_limit", "scattering_order_decomposition",
    "voigt_tmatrix_from_resonance_result", "sub_cell_centres", "strain_from_displacement_traction",
    "ResonanceOceanBottomResult", "compute_resonance_slab_scattering", "resonance_ocean_bottom",
    "resonance_slab_tmatrices", "uniform_contrast_phases", "subcell_slab",
    "compute_subcell_slab_scattering",
    "CPAResult", "CubicEffectiveMedium", "Phase", "compute_cpa", "compute_cpa_two_phase",
    "phases_from_two_phase", "voigt_average", "SphereDecompositionResult", "compute_sphere_foldy_lax",
    "compute_sphere_foldy_lax_fft", "compute_sphere_foldy_lax_fft_gpu",
]
