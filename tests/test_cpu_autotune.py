"""The autotuner's coordinate descent (host logic, no GPU): it finds the
separable optimum, never measures a configuration twice, keeps the defaults
when nothing is faster, and reuses a cached result."""
import json

from arapgs_amd import autotune as A


def _cost(cfg):
    # separable cost with its minimum away from the defaults
    return (abs(cfg["aggregates"] - 192) / 64 + abs(cfg["refresh"] - 16) / 4 + abs(cfg["host_checks"] - 8) / 8
            + (0.5 if cfg["tile"] == 8 else 0.0) + abs(cfg["blend_unroll"] - 8) / 4 + (0.3 if cfg["warm"] else 0.0)
            + abs(cfg["cg_block"] - 128) / 128 + (0.2 if cfg["cg_graphs"] == 0 else 0.0)
            + (0.1 if cfg["node_cache"] == 0 else 0.0) + abs(cfg["blend_split"] - 2048) / 8192
            + abs(cfg["deform_skip"] - 32) / 64 + (0.15 if cfg["cg_mode"] == 0 else 0.0)
            + (0.05 if cfg["cg_predict"] == 0 else 0.0) + abs(cfg["solve_cus"] - 64) / 256
            + (0.02 if cfg["samples_cus"] != 1 else 0.0) + (0.03 if cfg["sh_compose"] == 0 else 0.0)
            + (0.04 if cfg["tile_cull"] == 0 else 0.0) + abs(cfg["coarse_lag"] - 2) / 10
            + (0.01 if cfg["side_cus"] == 0 else 0.0) + (0.01 if cfg["samples_late"] == 0 else 0.0)
            + (0.01 if cfg["cg_asm"] == 0 else 0.0) + (0.01 if cfg["split_render"] == 0 else 0.0)
            + (0.01 if cfg["host_spin"] == 0 else 0.0) + 1.0)


def test_finds_the_separable_optimum():
    seen = []

    def measure(cfg):
        seen.append(tuple(sorted(cfg.items())))
        return _cost(cfg)

    best, trials = A.tune(measure)
    assert best == dict(tile=16, blend_unroll=8, aggregates=192, refresh=16, host_checks=8, warm=0, cg_block=128,
                        cg_graphs=1, cg_mode=2, cg_predict=1, node_cache=1, blend_split=2048, deform_skip=32,
                        solve_cus=64, samples_cus=1, sh_compose=1,
                        tile_cull=1, coarse_lag=2, side_cus=1,
                        samples_late=1, cg_asm=1,
                        split_render=1, host_spin=1)
    assert len(seen) == len(set(seen)) == len(trials)
    assert min(t for _, t in trials) == _cost(best)


def test_keeps_defaults_when_nothing_is_faster():
    best, _ = A.tune(lambda cfg: 1.0 + (0.0 if cfg == A.DEFAULTS else 0.1))
    assert best == A.DEFAULTS


def test_cache_is_reused(tmp_path):
    path = str(tmp_path / "tune.json")
    best, trials = A.tune(_cost, key="w1", cache_path=path)
    assert trials
    assert json.load(open(path))["w1"] == best
    best2, trials2 = A.tune(lambda cfg: 1 / 0, key="w1", cache_path=path)  # never measured
    assert best2 == best and trials2 == []
