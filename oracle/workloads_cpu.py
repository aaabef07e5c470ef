"""The benchmark workloads built on the CPU oracle (test infrastructure only:
used by tests/ and bench.py's cpu_baseline leg)."""
from __future__ import annotations


def config_oracle(O, cfg: dict):
    """Lattice::new + plate + init at cfg['init_u'], then the inflow set to
    cfg['inflow'] (the public field, src/lattice.rs:18)."""
    lat = O.Lattice(cfg["w"], cfg["h"], cfg["density"], cfg["init_u"], cfg["inlet"], cfg["rest"],
                    cfg["bands"], cfg["seed"])
    if cfg.get("plate"):
        lat.add_solid_box(*cfg["plate"])
    if cfg.get("profile") == "shear":
        import numpy as np

        from workloads import shear_profile  # the same f32 profile the engine gets
        uy = shear_profile(cfg["w"])
        lat.init_profile(0, np.zeros_like(uy), uy)
    else:
        lat.init_equilibrium()
    lat.set_inflow(*cfg["inflow"])
    return lat
