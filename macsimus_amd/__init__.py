"""MI355X-native force engine for MACSIMUS `cook` (see include/mdx.h)."""
