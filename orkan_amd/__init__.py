"""orkan_amd -- MI355X-native state-update path of the Orkan simulator."""
