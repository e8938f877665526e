"""MI355X-native render pass of hypertrace: scene vocabulary, codegen and the
HIP runtime binding. The compute path is the HIP library under lib/; nothing
here falls back to a CPU renderer."""
from . import objects, types  # noqa: F401

__all__ = ["objects", "types"]
