"""emc.conf parsing (configs.cpp:40-90 key semantics) through the C ABI; no GPU."""
import pytest

import emc_amd
from emc_amd import config

TEXT = """
# detector and model
model_side = 32; # a trailing comment
read_stride = 2;
wavelength = 2.5e-10;
pixel_size = 1.5e-4;
detector_size = 512;
detector_distance = 0.5;
rotations_n = 5;
sigma_start = 0.5;   sigma_final = 0.1;
sigma_half_life = 10;
diff_type = "absolute";
slice_chunk = 700;
N_images = 123;
max_iterations = 7;
image_prefix = "/data/run" "_"; // adjacent strings join
mask_file = "/data/mask.h5";
normalize_images = false;
known_intensity = TRUE;
model_input = 0;
model_file = "";
exclude_images = false;
output_period = 3;
output_loop = 1;
isOutput =1;
isDebug = 0;
/* engine keys */
precision = "fp32"; interp : "nearest"; insert = "atomic";
fluence = 2;
proposal_coarse_n = 2; proposal_top_k = 6;
"""


def test_parse_reference_keys_and_engine_keys():
    c = config.parse(TEXT)
    assert (c["model_side"], c["read_stride"], c["detector_size"], c["rotations_n"]) == (32, 2, 512, 5)
    assert c["wavelength"] == 2.5e-10 and c["pixel_size"] == 1.5e-4 and c["detector_distance"] == 0.5
    assert (c["sigma_start"], c["sigma_final"], c["sigma_half_life"]) == (0.5, 0.1, 10)
    assert c["diff"] == 0 and c["slice_chunk"] == 700 and c["N_images"] == 123 and c["max_iterations"] == 7
    assert c["image_prefix"] == "/data/run_" and c["mask_file"] == "/data/mask.h5"
    assert c["normalize_images"] == 0 and c["known_intensity"] == 1
    assert (c["output_period"], c["output_loop"], c["isOutput"], c["isDebug"]) == (3, 1, 1, 0)
    assert (c["precision"], c["interp"], c["insert"], c["fluence"]) == (0, 0, 0, 2.0)
    assert (c["proposal_coarse_n"], c["proposal_top_k"]) == (2, 6)


@pytest.mark.parametrize("kind,code", [("absolute", 0), ("poisson", 1), ("relative", 2), ("true_poisson", 3)])
def test_diff_types(kind, code):
    assert config.parse(f'diff_type = "{kind}";')["diff"] == code


def test_defaults_for_missing_keys():
    c = config.parse("")
    d = config.defaults()
    assert c == d
    assert d["model_side"] == 64 and d["diff"] == 3 and d["known_intensity"] == 1


@pytest.mark.parametrize("bad", ['model_sidee = 64;', 'model_side = 64.5;', 'diff_type = "chi";',
                                 'model_side 64;', 'image_prefix = "unterminated;', 'precision = "fp16";',
                                 'model_side = 2;', 'proposal_top_k = 0;', 'proposal_coarse_n = -1;'])
def test_errors_are_reported(bad):
    with pytest.raises(emc_amd.EmcError) as e:
        config.parse(bad)
    assert "line" in str(e.value) or "range" in str(e.value)


def test_read_file(tmp_path):
    p = tmp_path / "emc.conf"
    p.write_text(TEXT)
    assert config.read(str(p))["N_images"] == 123
    with pytest.raises(emc_amd.EmcError):
        config.read(str(tmp_path / "missing.conf"))
