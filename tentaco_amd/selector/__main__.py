from tentaco_amd.selector import main

main()
