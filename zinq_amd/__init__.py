"""zinq_amd: MI355X-native AO->MO ERI transform and J/K Fock build for Zinq's HF / MP / CI methods."""
