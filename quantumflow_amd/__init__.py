"""quantumflow_amd: QuantumFlow's state-vector API on AMD MI355X (gfx950).

The classes and functions mirror the reference package (qf.State, qf.Gate,
qf.Circuit, qf.zero_state, the standard gates, qf.qaoa ...) with the
backend replaced by hand-written HIP kernels behind a C-ABI (include/qfx.h).

    import quantumflow_amd as qf
    ket = qf.Circuit([qf.H(0), qf.CNOT(0, 1)]).run()
"""

from .config import version as __version__  # noqa: F401
from . import backend  # noqa: F401
from .qubits import *  # noqa: F401,F403
from .states import *  # noqa: F401,F403
from .ops import *  # noqa: F401,F403
from .gates import *  # noqa: F401,F403
from .stdgates import *  # noqa: F401,F403
from .stdops import *  # noqa: F401,F403
from .cbits import *  # noqa: F401,F403
from .programs import *  # noqa: F401,F403
from .circuits import *  # noqa: F401,F403
from .channels import *  # noqa: F401,F403
from .dagcircuit import *  # noqa: F401,F403
from .measures import *  # noqa: F401,F403
from . import qaoa  # noqa: F401
from . import utils  # noqa: F401
from . import forest  # noqa: F401
