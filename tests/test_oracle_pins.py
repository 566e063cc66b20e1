"""Pins of the oracle where textbook local SW / NW and the reference differ
(tests/oracle_pins.py), and a mutation run: plausible mistakes planted in a
copy of the oracle's C source, on the CPU, must each be caught by the pins."""
import os
import shutil
import subprocess
import sys

import pytest

from tests import oracle_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.parametrize("pin", P.LOCAL_PINS + P.NW_PINS, ids=lambda f: f.__name__)
def test_pin(pin):
    fails = pin()
    assert not fails, fails


# (name, source text, replacement) -- each a mistake a transcription of the
# reference could plausibly make
MUTANTS = [
    ("full lazy-F (textbook Farrar: repeat until nothing changes)",
     "        u8x16 cf = shift_right(f);                             /* :234-242 lazy-F */\n"
     "        for (int s = 0; s < sn; ++s) {",
     "        u8x16 cf = shift_right(f);\n"
     "        for (int pass = 0; pass < 16; ++pass, cf = shift_right(cf))\n"
     "        for (int s = 0; s < sn; ++s) {"),
    ("lazy-F without the early stop",
     "            if (same) break;",
     "            (void)same;"),
    ("best cell by query order (lane-major ties)",
     "                for (int l = 0; l < lanes; ++l) {\n"
     "                    if (H[s].v[l] == lm) {\n"
     "                        int j = s + l * sn;\n"
     "                        if (j < qlen) { max_j = j; break; }\n"
     "                    }\n"
     "                }",
     "                for (int jj = 0; jj < qlen; ++jj)\n"
     "                    if (H[jj % sn].v[jj / sn] == lm) { max_j = jj; break; }"),
    ("stripe maximum over real lanes only",
     "            for (int l = 0; l < lanes; ++l) lm = u8max(lm, H[s].v[l]);",
     "            for (int l = 0; l < lanes; ++l) if (s + l * sn < qlen) lm = u8max(lm, H[s].v[l]);"),
    ("score2 exclusion zone one wider",
     "if ((col < lo || col > hi) && s > score2) score2 = s;",
     "if ((col <= lo || col >= hi) && s > score2) score2 = s;"),
    ("NW deletion opens on ties",
     "uint8_t e_flag = e_open > e_ext ? 0 : 0x01;",
     "uint8_t e_flag = e_open >= e_ext ? 0 : 0x01;"),
    ("NW insertion opens on ties",
     "uint8_t f_flag = f_open > f_ext ? 0 : 0x02;",
     "uint8_t f_flag = f_open >= f_ext ? 0 : 0x02;"),
    ("NW H prefers E on ties",
     "uint8_t from_e = ev > h_cur ? 0x04 : 0;",
     "uint8_t from_e = ev >= h_cur ? 0x04 : 0;"),
    ("NW H prefers F on ties",
     "uint8_t from_f = fv > h_cur ? 0x08 : 0;",
     "uint8_t from_f = fv >= h_cur ? 0x08 : 0;"),
]


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(tmp_path, name, old, new):
    src = open(os.path.join(ORACLE, "sw_oracle.c")).read()
    assert src.count(old) == 1, "mutation site not found once: %s" % name
    for f in ("sw_oracle.h", "e2e_oracle.c", "e2e_oracle.h"):
        shutil.copy(os.path.join(ORACLE, f), tmp_path / f)
    (tmp_path / "sw_oracle.c").write_text(src.replace(old, new))
    lib = tmp_path / "libswor_mutant.so"
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-o", str(lib),
                    str(tmp_path / "sw_oracle.c"), str(tmp_path / "e2e_oracle.c"), "-lpthread", "-lm"],
                   check=True)
    r = subprocess.run([sys.executable, "-m", "tests.oracle_pins"], cwd=ROOT, capture_output=True,
                       text=True, env=dict(os.environ, ORACLE_LIB=str(lib)), timeout=300)
    last = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
    assert last.startswith("FAILURES"), r.stdout[-2000:] + r.stderr[-2000:]
    assert int(last.split()[1]) > 0, "the pins did not catch: %s" % name
