"""Copy the reference's ED test data (test/test_data of LadderDGA.jl) into this folder.

The three files are data the reference's own tests hold, kept verbatim:
  b1u2_ED_data.jld2  DMFT/ED inputs of the b1u2 system (U=2, β=1, half filling):
                     local vertices Γsp/Γch, generalized susceptibilities χDMFTsp/ch,
                     gImp, g0, bath ϵₖ/Vₖ, parameters and frequency-grid sizes
                     (read by test/full_run.jl through config_b1u2.toml and by
                     test/thermodynamics.jl:2)
  ed_out.csv         the ED kinetic energy and double occupancy of the same system
                     (the golden values of test/thermodynamics.jl:1-5)
  config_b1u2.toml   the reference's run configuration for it
Run once in a checkout that has the reference at /root/reference; tests only read the copies.
"""
import shutil
from pathlib import Path

SRC = Path("/root/reference/test/test_data")
DST = Path(__file__).resolve().parent
for name in ("b1u2_ED_data.jld2", "ed_out.csv", "config_b1u2.toml"):
    shutil.copyfile(SRC / name, DST / name)
    print("copied", name)
