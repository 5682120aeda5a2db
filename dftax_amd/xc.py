"""Exchange-correlation functional values (choices as values, as in dftax).

Each class names a functional the device kernels implement; ``hf_coeff`` is
the exact-exchange fraction (dftax/energy/xc.py:527-601). The formulas
themselves live in the HIP kernel (csrc/xc.hip).
"""

from __future__ import annotations


class XCFunctional:
    name: str = ""
    xc_type: str = "GGA"
    hf_coeff: float = 0.0
    # range-separated hybrids: K_hf = hf_coeff K + hf_coeff_lr K_lr(omega), K_lr
    # from erf(omega r)/r integrals (dftax/energy/xc.py:49-55); 0 for global hybrids
    hf_coeff_lr: float = 0.0
    omega: float = 0.0
    xc_id: int = 0

    def __repr__(self) -> str:
        return f"{self.name}()"

    def __eq__(self, other) -> bool:
        return type(self) is type(other)

    def __hash__(self) -> int:
        return hash(self.name)


class PBE(XCFunctional):
    name, hf_coeff, xc_id = "PBE", 0.0, 1


class PBE0(XCFunctional):
    """0.75 PBE exchange + PBE correlation + 25% exact exchange (xc.py:584-601)."""
    name, hf_coeff, xc_id = "PBE0", 0.25, 2


class B3LYP(XCFunctional):
    """0.08 Slater + 0.72 B88 + 0.19 VWN_RPA + 0.81 LYP + 20% exact exchange (xc.py:527-567)."""
    name, hf_coeff, xc_id = "B3LYP", 0.20, 3


class CAMB3LYP(XCFunctional):
    """CAM-B3LYP (xc.py:711-750): 0.35 B88 + 0.46 short-range B88 (ITYH, omega
    0.33) + 0.81 LYP + 0.19 VWN5, exact exchange 0.19 K + 0.46 K_lr(0.33)."""
    name, hf_coeff, hf_coeff_lr, omega, xc_id = "CAM-B3LYP", 0.19, 0.46, 0.33, 4


def resolve(xc) -> XCFunctional:
    if isinstance(xc, XCFunctional):
        return xc
    if isinstance(xc, type) and issubclass(xc, XCFunctional):
        return xc()
    table = {"PBE": PBE, "PBE0": PBE0, "B3LYP": B3LYP, "CAM-B3LYP": CAMB3LYP, "CAMB3LYP": CAMB3LYP}
    if isinstance(xc, str) and xc.upper() in table:
        return table[xc.upper()]()
    raise ValueError(f"unsupported functional {xc!r}; available: PBE, PBE0, B3LYP, CAM-B3LYP")
