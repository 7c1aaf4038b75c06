"""MI355X-native plane-wave EPM band solver (H(k) assembly + Hermitian eigensolve)."""
from .material import Material, Materials, load as load_material  # noqa: F401
