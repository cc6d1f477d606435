"""On-device alpha selectors against the oracle (GPU).

Features and alpha of pgj_alpha_select for seeded filtered-noise designs of
several contrasts and a checkerboard, against oracle/select.py: log10
contrast and density moments to 1e-13, the FP32 law features to 1e-6, the
1/16-rounded selectors exactly, compact8 to 1e-6. A resident solve with a
device-selected alpha (pgj_resident_scale_dev) equals the solve with the same
alpha given as a number.
"""
import numpy as np
import pytest
import torch

import pgj
from pgj.select import FEATURES, Selector
from oracle import design
from oracle import select as sel

pytestmark = pytest.mark.gpu


def fields():
    out = []
    for n, emin, radius in ((16, 1e-4, 2), (24, 1e-6, 2), (16, 1e-9, 1)):
        rho = design.filtered_noise(n, 20261104 + n, radius)
        out.append((f'noise{n}_{emin:g}', design.simp(rho, emin), rho))
    n = 16
    z, y, x = np.indices((n, n, n))
    cb = np.where((x + y + z) % 2 == 0, 1e-9, 1.0)
    out.append(('checkerboard', cb, np.where(cb < 1, 0.0, 1.0)))
    return out


@pytest.mark.parametrize('name,a,rho', fields(), ids=[f[0] for f in fields()])
def test_selectors_match_oracle(name, a, rho):
    print(f'selectors {name}', flush=True)
    op = pgj.Operator(torch.tensor(a, device='cuda'))
    s = Selector(op.grid)
    at, rt = torch.tensor(a, device='cuda'), torch.tensor(rho, device='cuda')
    got = {'gray': float(s(at, 'gray', rho=rt).item())}
    moments = s.features.cpu().numpy()[1:3]          # the density moments come with the gray selector
    got['compact8'] = float(s(at, 'compact8').item())
    feats = dict(zip(FEATURES, s.features.cpu().numpy()))   # the law features come with the compact selectors
    feats['mean_rho'], feats['mean_rho2'] = moments
    for m in ('contrast', 'compact8q', 'green', 'gj', 0.3125):
        got[m] = float(s(at, m).item())
    assert abs(feats['log10_contrast'] - sel.log10_contrast(a)) < 1e-13
    assert abs(feats['mean_rho'] - rho.mean()) < 1e-13 and abs(feats['mean_rho2'] - np.square(rho).mean()) < 1e-13
    area, jump, contrast, complexity = sel.law_features(a)
    for key, ref in (('area', area), ('jump', jump), ('law_contrast', contrast), ('complexity', complexity)):
        assert abs(feats[key] - ref) <= 1e-6 * max(abs(ref), 1e-12), (key, feats[key], ref)
    assert got['contrast'] == sel.contrast_alpha(a)
    assert got['gray'] == sel.gray_alpha(a, rho)
    assert got['compact8q'] == sel.compact8_alpha(a, quantize=True)
    assert abs(got['compact8'] - sel.compact8_alpha(a)) < 1e-6
    assert (got['green'], got['gj'], got[0.3125]) == (0.0, 1.0, 0.3125)
    assert float(s(at, 'eta', emin=1e-6).item()) == sel.eta_alpha(1e-6)


def test_device_alpha_solve_equals_numeric_alpha():
    n = 16
    a = design.simp(design.filtered_noise(n, 7, 2), 1e-6)
    at = torch.tensor(a, device='cuda')
    dev = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    dev.material(at)
    alpha = float(dev.select('contrast').item())
    assert alpha == sel.contrast_alpha(a) and alpha > 0
    i_dev, _ = dev.solve(rtol=1e-10)
    num = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    num.material(at)
    num.select(alpha)
    torch.testing.assert_close(dev.scale, num.scale, rtol=0, atol=0)
    i_num, _ = num.solve(rtol=1e-10)
    assert [i['iterations'] for i in i_dev] == [i['iterations'] for i in i_num]
    torch.testing.assert_close(dev.effective_tensor(), num.effective_tensor(), rtol=0, atol=0)
    with pytest.raises(ValueError):
        dev.select('gray')                # needs the physical density
