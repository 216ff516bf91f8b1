"""MI355X-native drop-in for the CCPi-Reconstruction projector + iterative
solver path (package layout mirrors the reference's ``ccpi.reconstruction``)."""
