"""analisi_amd: the time-lagged pair-correlation path of rikigigi/analisi
(Gofrt g(r,t), mean square displacement, Green-Kubo lag correlations and their
block averages) on an MI355X (gfx950).

``analisi_amd.pyanalisi`` has the pyanalisi classes of this path (Trajectory,
ReadLog, BoxFormat, Gofrt, MeanSquareDisplacement, GreenKubo) with the
reference's constructor arguments and buffer layouts; ``analisi_amd.capi`` is
the ctypes binding of the C ABI (include/analisi_hip.h).  Both load the in-tree
libanalisi_hip.so; there is no CPU fallback.
"""
import os

__all__ = ["capi", "pyanalisi", "synth", "build"]
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
