"""Seeded synthetic scenes of the benchmark workloads and the tests: the
20-band clustered fields (``field``), the independent island cutouts
(``cutouts``) and the serving stream's tiles (``tiles``).  Data generation
only; the fit path (``uniphot_amd``) never imports it."""
