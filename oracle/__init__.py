"""FP64 CPU oracle for the forced-photometry fit path (TEST INFRASTRUCTURE).

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``uniphot_amd``) never imports, calls or falls
back to anything here; it fails loudly when its HIP library is missing.

Every function restates the reference algorithm in float64 NumPy and cites
the reference file:line it follows (paths relative to the uniphot checkout,
package ``tract7dt``).

Parity pins (what checks each oracle function independently of its own
formula) are listed in each module header.  Functions with no independent
pin say "parity unpinned" there.
"""
