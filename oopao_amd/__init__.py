"""oopao_amd: MI355X-native drop-in for OOPAO's optical-propagation hot path.

The classes mirror OOPAO's Telescope, Source, Atmosphere, DeformableMirror,
Detector and Pyramid (constructor signatures, tags, operators ``src**atm*tel*dm*wfs``,
properties), with the per-frame numerics on the GPU through the HIP library
``liboopao_amd.so`` (C-ABI in include/oopao_amd.h). Module paths follow the
reference (``from oopao_amd.Pyramid import Pyramid``).
"""
from .tools import OopaoError, warning  # noqa: F401
from .Source import Source  # noqa: F401
from .Telescope import Telescope  # noqa: F401
from .Detector import Detector  # noqa: F401
from .MisRegistration import MisRegistration  # noqa: F401
from .DeformableMirror import DeformableMirror  # noqa: F401
from .Pyramid import Pyramid  # noqa: F401
from .GainSensingCamera import GainSensingCamera  # noqa: F401
from .BioEdge import BioEdge  # noqa: F401
from .Asterism import Asterism  # noqa: F401
from .Atmosphere import Atmosphere  # noqa: F401
from .calibration.InteractionMatrix import InteractionMatrix, InteractionMatrixFromPhaseScreen  # noqa: F401
from .calibration.CalibrationVault import CalibrationVault  # noqa: F401
