// placeholder: multithreaded CPU LSM baseline (filled in below)
