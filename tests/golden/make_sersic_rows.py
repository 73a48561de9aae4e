"""Write tests/golden/sersic_rows.npz: a sample of the reference's Sersic
mixture table rows (data file tract7dt/data/sersic_mog_table.npz, read with
numpy.load, allow_pickle=False).  Rows n >= 0.8 (Shajib family) pin the
table generator; a few n < 0.8 rows are kept to document the difference.
Run in the container that holds the reference checkout."""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tract7dt/data/sersic_mog_table.npz")
OUT = Path(__file__).with_name("sersic_rows.npz")


def main():
    z = np.load(SRC, allow_pickle=False)
    n = z["n_knots"]
    shajib = np.arange(n.size - 1101, n.size)          # the n >= 0.8 segment
    pick = np.unique(np.concatenate([shajib[::50], shajib[-3:], shajib[:3]]))
    low = np.arange(0, n.size - 1101, 40)
    np.savez(OUT, idx=pick, n=n[pick], amps=z["amps"][pick], log_vars=z["log_vars"][pick],
             core=z["core_amp"][pick], low_n=n[low], low_amps=z["amps"][low],
             low_log_vars=z["log_vars"][low], low_core=z["core_amp"][low])
    print(OUT, pick.size, low.size)


if __name__ == "__main__":
    main()
