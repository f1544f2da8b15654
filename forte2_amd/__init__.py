"""forte2_amd -- MI355X-native density-fitted J/K builder for forte2's SCF loops.

The hot path (J/K from the HBM-resident B(P|mn)) runs in hand-written gfx950
HIP kernels behind the C ABI of include/f2jk.h; ``FockBuilder`` is the drop-in
for ``forte2.jkbuilder.FockBuilder``.
"""

__version__ = "0.1.0"


def __getattr__(name):
    # lazy imports keep `import forte2_amd` free of torch/HIP initialisation
    if name == "FockBuilder":
        from .jkbuilder import FockBuilder

        return FockBuilder
    raise AttributeError(name)
