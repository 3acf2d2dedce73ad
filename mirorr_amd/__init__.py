"""MI355X-native block-matching registration (Mirorr hot path)."""
