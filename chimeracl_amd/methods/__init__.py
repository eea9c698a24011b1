from .generic_methods_hip import Communicator, GenericMethodsHIP, default_comm  # noqa: F401
from .grid_methods_hip import GridMethodsHIP  # noqa: F401
from .particles_methods_hip import ParticleMethodsHIP  # noqa: F401
from .solver_methods_hip import SolverMethodsHIP  # noqa: F401
from .transformer_methods_hip import TransformerMethodsHIP  # noqa: F401
