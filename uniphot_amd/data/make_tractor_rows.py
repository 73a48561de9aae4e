"""Write sersic_tractor_rows.npz: the Sersic mixture rows below the splice
index n = 0.8 of the reference's table (tract7dt/data/sersic_mog_table.npz,
read with numpy.load, allow_pickle=False) -- every knot of its Tractor
segments, the duplicated segment-boundary knots and the segment's closing
n = 0.8 row included.  Those rows are numeric fits produced by The Tractor's
SersicMixture and exist only as data; the rows from n = 0.8 up are
regenerated by this package (sersic_table.py).  Run once in a checkout that
holds the reference; the output is committed."""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tract7dt/data/sersic_mog_table.npz")
OUT = Path(__file__).with_name("sersic_tractor_rows.npz")


def main():
    z = np.load(SRC, allow_pickle=False)
    n = z["n_knots"]
    splice = float(z["splice_n"])
    # the Tractor family ends at the first knot equal to the splice index
    end = int(np.searchsorted(n, splice, side="left")) + 1
    assert n[end - 1] == splice and n[end] == splice
    np.savez(OUT, n_knots=n[:end], amps=z["amps"][:end], log_vars=z["log_vars"][:end],
             core_amp=z["core_amp"][:end], n_min=float(z["n_min"]), splice_n=splice)
    print(OUT, end)


if __name__ == "__main__":
    main()
