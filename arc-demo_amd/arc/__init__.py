"""arc: the MI355X engine's Python package (ctypes layer over libarcsim)."""
