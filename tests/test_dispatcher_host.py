"""The ionic plugin aggregators on the host (no GPU): region maps
(hand-worked), validate_region_map's warning, SilentCell through the
RegionIonicModel protocol, and NodeIonicDispatcher's group logic with models
that are not the HIP-backed phenotypes (a user's torch model, SilentCell),
mirroring hrdayapy/simulation/functions/ionic_models/dispatcher.py:140-209."""
import numpy as np
import pytest
import torch

from hrdayapy_amd.simulation.functions.ionic_models import (
    LABEL_ENDO, LABEL_EPI, LABEL_MCELL, NodeIonicDispatcher, RegionIonicModel, SilentCell,
    make_transmural_region_map, region_map_from_segmentation, validate_region_map)


class Leak:
    """A user model: I = g (V - E), one state variable counting its calls."""
    name = "Leak"

    def __init__(self, g, E):
        self.g, self.E = g, E

    def state_names(self):
        return ["calls"]

    def init_state_tensor(self, N, device, dtype):
        return torch.zeros((N, 1), dtype=dtype, device=device)

    def step(self, V, state, dt, I_stim):
        return self.g * (V - self.E), state + 1.0


def test_region_maps():
    seg = np.zeros((3, 3, 2), dtype=np.int32)
    seg[1, :, :] = 3
    seg[2, 2, 1] = 7
    vox = np.argwhere(np.ones((3, 3, 2), dtype=bool))
    rm = region_map_from_segmentation(seg, vox)
    assert rm.dtype == np.int32 and rm.tolist() == [seg[tuple(v)] for v in vox]
    phi = np.array([0.0, 0.349, 0.35, 0.649, 0.65, 1.0, np.nan])
    assert make_transmural_region_map(phi).tolist() == [
        LABEL_ENDO, LABEL_ENDO, LABEL_MCELL, LABEL_MCELL, LABEL_EPI, LABEL_EPI, LABEL_MCELL]
    with pytest.warns(UserWarning, match=r"\[0, 7\]"):
        validate_region_map(np.array([0, 1, 7, 1]), {1: SilentCell()})


def test_silent_cell_protocol():
    m = SilentCell()
    assert isinstance(m, RegionIonicModel) and m.state_names() == []
    st = m.init_state_tensor(5, torch.device("cpu"), torch.float64)
    assert st.shape == (5, 0)
    I, st2 = m.step(torch.ones(5, dtype=torch.float64), st, 0.02, 0.0)
    assert torch.equal(I, torch.zeros(5, dtype=torch.float64)) and st2 is st


def test_dispatcher_groups_without_gpu_models():
    rm = np.array([2, 0, 2, 5, 0, 2])
    disp = NodeIonicDispatcher(rm, {2: Leak(0.5, -80.0), 5: Leak(2.0, 0.0)}, "cpu",
                               torch.float64)
    assert disp.region_labels() == [0, 2, 5]
    assert isinstance(disp.model_for_region(0), SilentCell)
    assert disp.node_indices_for_region(2).tolist() == [0, 2, 5]
    V = torch.tensor([-70.0, -10.0, -90.0, 10.0, 5.0, -80.0], dtype=torch.float64)
    I = disp.step(V, 0.02)
    np.testing.assert_allclose(I.numpy(), [5.0, 0.0, -5.0, 20.0, 0.0, 0.0])
    disp.step(V, 0.02, I_stim=torch.zeros(6, dtype=torch.float64))
    assert disp.get_state(2)[:, 0].tolist() == [2.0, 2.0, 2.0]
    disp.set_state(5, torch.full((1, 1), 7.0, dtype=torch.float64))
    assert disp.get_state(5).item() == 7.0
    with pytest.raises(KeyError):
        disp.model_for_region(3)
