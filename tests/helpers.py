"""Shared helpers of the test suite."""
from evora_amd.params import HydroParams


def params_from_index(entry):
    return HydroParams(**entry["params"])
