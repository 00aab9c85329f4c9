"""CorrectorNet structure on the CPU: output shape, the zero head of the
reference (every correction zero until trained), the seeded head, and the
module / parameter names of src/neuralpom/models/corrector_net.py, so that a
state_dict of the reference's network loads into this one."""
import torch

from neuralpom_amd.hybrid import CorrectorNet


def test_shapes_and_heads():
    x = torch.randn(1, 101, 26, 22, generator=torch.Generator().manual_seed(0))
    net = CorrectorNet(hidden_dim=8)
    y = net(x)
    assert y.shape == (1, 60, 26, 22)
    assert float(y.abs().max()) == 0.0
    seeded = CorrectorNet(hidden_dim=8, head_std=0.05, generator=torch.Generator().manual_seed(1))
    assert float(seeded(x).abs().max()) > 0.0


def test_parameter_names_follow_the_reference_layout():
    net = CorrectorNet(hidden_dim=64)
    sd = net.state_dict()
    for key, shape in (("inc.double_conv.0.weight", (64, 101, 3, 3)),
                       ("inc.double_conv.1.weight", (64,)),
                       ("inc.double_conv.3.weight", (64, 64, 3, 3)),
                       ("inc.shortcut.0.weight", (64, 101, 1, 1)),
                       ("down1.maxpool_conv.1.double_conv.0.weight", (128, 64, 3, 3)),
                       ("down3.maxpool_conv.1.shortcut.1.bias", (512,)),
                       ("up1.conv.double_conv.0.weight", (384, 768, 3, 3)),
                       ("up1.conv.double_conv.3.weight", (256, 384, 3, 3)),
                       ("up3.conv.shortcut.0.weight", (64, 192, 1, 1)),
                       ("outc.weight", (60, 64, 1, 1)), ("outc.bias", (60,))):
        assert tuple(sd[key].shape) == shape, key
    # GroupNorm groups: min(32, c) reduced to a divisor of c
    assert net.inc.double_conv[1].num_groups == 32
    assert CorrectorNet(hidden_dim=12).inc.double_conv[1].num_groups == 12
    assert CorrectorNet(hidden_dim=40).inc.double_conv[1].num_groups == 20
