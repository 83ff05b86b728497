"""MI355X-native mechanical equilibration (MEDYAN path): C-ABI library bindings."""
