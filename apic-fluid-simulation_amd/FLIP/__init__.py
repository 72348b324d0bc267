"""FLIP -- MI355X-native APIC fluid step behind the reference's `FLIP::apic` API.

`FLIP.apic` mirrors the Rust module of the reference (src/lib.rs:1,
src/apic.rs): `Apic`, `WorldProperties`, `DebugOptions`.  Every call goes
through the C ABI of libapic_hip.so (include/apic.h); there is no CPU path.
"""
from . import apic, scenes  # noqa: F401
