"""diffren_amd: MI355X-native rasterize -> interpolate -> splat for diffren."""
__version__ = "0.1.0"
