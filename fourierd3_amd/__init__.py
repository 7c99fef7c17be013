"""MI355X (gfx950) implementation of FourierD3's D3(BJ) energy / force / stress path.

Exports the reference's two public names (``src/fourierd3/__init__.py:1-7``).
"""
from fourierd3_amd.core import FourierD3
from fourierd3_amd.calculator import FourierD3ASECalculator

__all__ = (
    "FourierD3",
    "FourierD3ASECalculator",
)
