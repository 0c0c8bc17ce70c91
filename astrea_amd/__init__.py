"""MI355X (gfx950) finite-volume update for astrea (mervyzr/astrea).

Host Python on PyTorch-ROCm tensors; the update itself runs in hand-written
HIP kernels reached through the C ABI of libastrea_hip.so
(include/astrea_hip.h).  ``options`` mirrors astrea's option surface,
``plan.Plan`` owns a device plan, ``core.core_run`` is the drop-in for the
reference's time loop (astrea.py:31-85).
"""
__all__ = ["options", "plan", "core"]
