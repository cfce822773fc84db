"""srmg: MI355X-native block-structured AMR geometric multigrid.

Host side of the library: block tree and connectivity (``tree``), host block
fields in the reference's layout (``blockfield``), seeded synthetic problems
(``configs``) and the ctypes binding of the HIP shared library (``solver``).
"""
