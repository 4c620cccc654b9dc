"""MI355X-native per-step contact pipeline behind a pypybox2d-compatible API."""
