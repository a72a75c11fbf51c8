"""CPU oracle for the neuralangelo_amd hot path — TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithm of every function on
the hot path (NVlabs/neuralangelo; each function cites the file:line it
follows).  It is the checker, never the thing measured or shipped: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, and the product package ``neuralangelo_amd`` never does.

Modules
  hashgrid  — level table, exact fmaf, corner indices/weights, encode (torch CPU)
  field     — SDF MLP, colour MLP, SH, 4-tap numerical gradient (torch CPU)
  render    — NeuS alphas, compositing, composite, sampling, losses, render_rays
  mesh      — lattice blocks + marching-cubes driver over the C restatement
  mc_oracle.c — marching cubes (PyMCubes algorithm) in plain C

Pinning status (what each restatement is checked against) is listed in the
docstring of each module; functions without an independent check say
"parity unpinned" there.
"""
