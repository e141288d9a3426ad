"""Drop-in descriptor classes (reference ``pyxtal_ff/descriptors``)."""
