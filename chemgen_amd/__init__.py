"""chemgen_amd: MI355X-native getrates (FP64 HIP for gfx950) behind Chemgen's interface."""
from .mechanism import Mechanism, load_mechanism, parse_chemkin, load_yaml  # noqa: F401

__all__ = ["Mechanism", "load_mechanism", "parse_chemkin", "load_yaml"]
