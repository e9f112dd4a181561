"""cctag_amd: MI355X-native CCTag detection hot path."""
