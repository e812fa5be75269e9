"""Backend availability (reference: hornlab_metal_bem/backends.py:9-89)."""
from __future__ import annotations

from dataclasses import dataclass


class AssemblyBackendUnavailable(RuntimeError):
    """Raised when the HIP library or a gfx950 device is not available."""


@dataclass(frozen=True)
class HipBackendStatus:
    library_path: str | None
    library_loaded: bool
    device_count: int
    reason: str | None

    @property
    def available(self) -> bool:
        return self.library_loaded and self.device_count > 0


def discover_hip_backend() -> HipBackendStatus:
    from .native import discover_native_runtime

    rt = discover_native_runtime()
    return HipBackendStatus(
        library_path=str(rt.library_path) if rt.library_path else None,
        library_loaded=rt.library_loaded,
        device_count=rt.device_count,
        reason="; ".join(rt.unavailable_reasons) or None,
    )


def resolve_assembly_backend() -> HipBackendStatus:
    status = discover_hip_backend()
    if not status.available:
        raise AssemblyBackendUnavailable(status.reason or "HIP backend is unavailable.")
    return status
