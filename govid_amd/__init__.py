"""govid_amd — MI355X-native EBIV decoder (gfx950, HIP) behind govid's Codec interface."""
