"""cmpy_amd: MI355X-native exact-diagonalisation engine for the Hubbard/Anderson
sector Hamiltonian, Lanczos ground state and Lanczos/Lehmann Green's functions.

Same names as cmpy's public surface for this path (``Basis``, ``Sector``,
``project_hopping``, ``HamiltonOperator``, ``HubbardModel``,
``SingleImpurityAndersonModel``, ``compute_groundstate``, ``gf_lehmann``, ...);
the compute runs in libcmpy_amd.so (HIP, gfx950) behind a C ABI.
"""

from ._lib import EDError, EDLibraryError, EDInvalidArgument, EDNumericError
from .basis import (
    UP, DN, SPIN_CHARS, state_label, binstr, binarr, binidx, overlap, occupations, create,
    annihilate, SpinState, State, Sector, Basis, upper_sector, lower_sector,
)
from .operators import (
    project_up, project_dn, project_elements_up, project_elements_dn, project_onsite_energy,
    project_hubbard_inter, project_hopping, LinearOperator, HamiltonOperator,
    SectorHamiltonOperator, CreationOperator, AnnihilationOperator,
)
from .models import (
    ModelParameters, AbstractModel, AbstractManyBodyModel, HubbardModel,
    SingleImpurityAndersonModel, chain_neighbors, square_neighbors,
)
from .exactdiag import (
    EigenState, groundstate, compute_groundstate, solve_sector, GreensFunctionMeasurement,
    gf_lehmann, greens_lehmann_t0, greens_cf, iter_lanczos_coeffs, lanczos_coeffs,
    lanczos_matrix, lanczos_ground_state, gf_greater, gf_lesser, gf_tevo, greens_lehmann, greens_cf_ij,
    occupation_up, occupation_dn, occupation, double_occupation, accumulate_gf, fourier_t2z,
)

from .streams import map_streams

__version__ = "0.1.0"
