"""CPU oracle for the tart-gnss-acquire acquisition search -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy, the reference algorithm of the path the HIP library
accelerates (``src/correlate.rs``, the per-constellation replica generators, ``acr.rs``,
``rfi.rs``, ``stats.rs`` and the JSON shape of ``lib.rs``/``acquisition.rs``).  It is the
checker the parity tests compare the GPU library against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it,
and only as the checker or as the timed CPU port; the product path never calls it.

Pinning (what independently checks each function here):
  codes      -- reference KAT vectors (tests/golden/code_kats.json), IS-GPS-200 first-10-chip
                table, Gold-code correlation sidelobes, Legendre residue count.
  correlate  -- brute-force O(N^2) circular correlation, tie rules on constructed inputs,
                the reference's delay-recovery identity, the two-period folding identity.
  acr        -- the reference's ACR known answers (r=4.1725 -> 45, 11.4497 -> 50 dB-Hz),
                exact table points, saturation, cut-off, monotone interpolation, and the
                log-space rule (r at the geometric mean of two brackets -> the arithmetic
                mean of their C/N0).
  rfi        -- closed forms: a tone on an FFT bin (|X_k|^2 = (A n / 2)^2 against a direct
                DFT), the floor index sorted[half/2], the >/>= local-maximum rule on
                plateaus, the strict 10 kHz clustering, the three-line cap and order, the
                (n - lag) ACF normalisation on periodic sequences, the population-variance
                dead-channel threshold; the reference's square-wave / white-noise /
                dead-channel tests.
  stats      -- the reference's median/MAD known answers (stats.rs:32-61).
  jsonfmt    -- shortest round-trip float formatting checked against float() round trip
                and hand-written serde_json/ryu expectations.
  antenna_test -- the eleven known-answer cases of antenna_test.rs's unit tests
                (tests/golden/antenna_test_kats.json: inputs and the asserted fields).
"""
