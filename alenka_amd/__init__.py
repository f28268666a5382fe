"""alenka_amd: MI355X-native relational operator hot path (filter -> decode -> join ->
group-by -> order/top-k) behind a C ABI (include/alenka_hip.h, libalenka_hip.so).

`alenka_amd.device` holds the torch-backed handles (needs a HIP device); `alenka_amd.host`
the operator layer that mirrors the reference's emit_* interface.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
