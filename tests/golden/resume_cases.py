"""Decks of the resume goldens (inputs only): each runs with intermediate
result files, and ``resume`` names the file a second run restarts from."""
from __future__ import annotations

from cases import CASES, _c1, _euler_1d, _shock_tube_1d

RESUME_CASES = {
    "resume_kapila_1d_o2": dict(_shock_tube_1d("kapila", 2, ["vanleer"] * 4, 200, cells=200), iter_freq=50, resume=2),
    "resume_multip_1d_o2_relax": dict(_shock_tube_1d("multip", 2, ["minmod"] * 4, 160, cells=100, relax=True),
                                      iter_freq=40, resume=3),
    "resume_euler_1d_fo_wall": dict(_euler_1d(1, ["none"] * 4, 300), bc=["abs", "wall"] + ["abs"] * 4,
                                    iter_freq=100, resume=1),
    "resume_kapila_1d_fo_time": dict(_c1(False, t_final=240e-6, cells=200), time_freq=60e-6, resume=2),
    "resume_kapila_2d_fo_injection": dict(CASES["kapila_2d_fo_injection"], iter_freq=20, resume=3),
}
