"""MaBoSS engines.  ``maboss.gpu_amd`` is the MI355X (gfx950) engine."""
