"""Save-step schedules (utils/rollout_schedules.py:20-90): known answers
worked out by hand from the schedules' definitions, and their use as
trajectory_rollout strides."""

from __future__ import annotations

import numpy as np
import pytest

from jaxdem_amd.utils.rollout_schedules import (
    make_save_steps_linear,
    make_save_steps_pseudolog,
    strides_from_steps,
)


def test_linear_known_answers():
    assert make_save_steps_linear(num_steps=10, save_freq=3).tolist() == [0, 3, 6, 9]
    assert make_save_steps_linear(num_steps=9, save_freq=3, include_step0=False).tolist() == [3, 6, 9]
    assert make_save_steps_linear(num_steps=0, save_freq=5).tolist() == [0]
    assert make_save_steps_linear(num_steps=2, save_freq=5, include_step0=False).tolist() == []
    with pytest.raises(ValueError):
        make_save_steps_linear(num_steps=-1, save_freq=1)
    with pytest.raises(ValueError):
        make_save_steps_linear(num_steps=10, save_freq=0)


def test_pseudolog_one_block_decades():
    # one block of 1000 steps, base spacing 1, decades of 10: every step up to
    # 10, every 10th up to 100, every 100th up to 1000
    got = make_save_steps_pseudolog(num_steps=1000, reset_save_decade=1000, min_save_decade=1)
    want = list(range(0, 11)) + list(range(20, 101, 10)) + list(range(200, 1001, 100))
    assert got.tolist() == want


def test_pseudolog_blocks_restart():
    # blocks of 10 steps, decades of 2 from spacing 1: offsets 1, 2 (spacing
    # 1), 4 (spacing 2 on (2, 4]), 8 (spacing 4 on (4, 8]); (8, 16] has
    # spacing 8 and no multiple of 8 in (8, 10]
    got = make_save_steps_pseudolog(num_steps=25, reset_save_decade=10, min_save_decade=1, decade=2)
    assert got.tolist() == [0, 1, 2, 4, 8, 11, 12, 14, 18, 21, 22, 24]
    no0 = make_save_steps_pseudolog(num_steps=25, reset_save_decade=10, min_save_decade=1, decade=2,
                                    include_step0=False)
    assert no0.tolist() == got.tolist()[1:]


def test_pseudolog_block_end_and_base_spacing():
    # spacing 5 up to 50, 50 up to 500: offsets 5..50 step 5, then 100..500
    # step 50 -- a block of 200 steps ends on its offset 200
    got = make_save_steps_pseudolog(num_steps=400, reset_save_decade=200, min_save_decade=5)
    first = list(range(0, 51, 5)) + [100, 150, 200]
    second = [200 + o for o in list(range(5, 51, 5)) + [100, 150, 200]]
    assert got.tolist() == first + second
    assert np.all(np.diff(got) > 0)


def test_pseudolog_cap_and_errors():
    full = make_save_steps_pseudolog(num_steps=1000, reset_save_decade=1000, min_save_decade=1)
    assert make_save_steps_pseudolog(num_steps=1000, reset_save_decade=1000, min_save_decade=1,
                                     cap=7).tolist() == full[:7].tolist()
    assert make_save_steps_pseudolog(num_steps=0, reset_save_decade=3, min_save_decade=1).tolist() == [0]
    for kw in (dict(num_steps=-1), dict(reset_save_decade=0), dict(min_save_decade=0), dict(decade=1)):
        args = dict(num_steps=10, reset_save_decade=5, min_save_decade=1, decade=10)
        args.update(kw)
        with pytest.raises(ValueError):
            make_save_steps_pseudolog(**args)


def test_strides_from_steps():
    steps = make_save_steps_pseudolog(num_steps=25, reset_save_decade=10, min_save_decade=1, decade=2)
    k = strides_from_steps(steps)
    assert k[0] == 0 and np.cumsum(k).tolist() == steps.tolist()
    with pytest.raises(ValueError):
        strides_from_steps([3, 2])


@pytest.mark.gpu
def test_pseudolog_frames_on_device(device):
    """A pseudo-log schedule drives the on-device frames: frame f is the
    state at step steps[f]."""
    import torch

    import jaxdem_amd as jd
    from jaxdem_amd import workloads as W

    print("[rollout] pseudolog frames", flush=True)
    w = W.spheres_friction(N=2048, seed=3, skin=0.05)
    st, sy = W.make_gpu_system(w, device=device)
    steps = make_save_steps_pseudolog(num_steps=60, reset_save_decade=30, min_save_decade=1, decade=3)
    _, _, traj = jd.System.trajectory_rollout(st, sy, strides=strides_from_steps(steps), save_fn=jd.save_state)
    assert traj.pos_c.shape[0] == len(steps)
    for f in (0, len(steps) // 2, len(steps) - 1):
        ref = jd.System.step(st, sy, n=int(steps[f]))[0] if steps[f] else st
        assert torch.equal(traj.pos_c[f], ref.pos_c), f
        assert torch.equal(traj.vel[f], ref.vel), f
