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
    lat.init_equilibrium()
    lat.set_inflow(*cfg["inflow"])
    return lat
