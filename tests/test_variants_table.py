"""The compiled variant table (csrc/spr_tuned.hpp) against the library, without
a GPU: every tuned entry names a compiled variant and is the default a new
handle of that size starts with; the tuner's tie rule and header writer."""
import os
import re

import numpy as np

from conftest import ROOT


def tuned_entries():
    text = open(os.path.join(ROOT, "s_spi_spring_amd", "csrc", "spr_tuned.hpp")).read()
    body = text[text.index("kTuned[]"):]
    return [tuple(map(int, m)) for m in re.findall(r"\{\s*(\d+),\s*(\d+),\s*(\d+)\s*\}", body)]


def test_tuned_table_names_compiled_variants():
    from s_spi_spring_amd import hipcore
    entries = [e for e in tuned_entries() if e[0] != 0]
    assert entries
    for n, col, row in entries:
        nc, nr = hipcore.variant_counts(n)
        assert 0 <= col < nc and 0 <= row < nr, (n, col, row, nc, nr)
        assert hipcore.default_variants(n) == (col, row)


def test_every_supported_size_has_variants():
    from s_spi_spring_amd import hipcore
    lib = hipcore.load()
    for n in (64, 128, 256, 384, 512, 768, 1024, 2048):
        assert lib.spr_supported_size(n) == 1
        nc, nr = hipcore.variant_counts(n)
        assert nc >= 1 and nr >= 1
    assert lib.spr_supported_size(640) == 0


def test_tie_rule_and_header():
    from s_spi_spring_amd import autotune
    assert autotune.pick(np.array([1.0, 0.5, 0.505])) == 1
    assert autotune.pick(np.array([0.507, 0.5, 0.9])) == 0       # within 1.5 %: the lower index
    assert autotune.pick(np.array([0.509, 0.5, 0.9])) == 1
    txt = autotune.header_text([{"n": 512, "column_variant": 1, "row_variant": 0},
                                {"n": 1024, "column_variant": 6, "row_variant": 1}], "test device")
    assert "{512, 1, 0}" in txt and "{1024, 6, 1}" in txt and "test device" in txt
